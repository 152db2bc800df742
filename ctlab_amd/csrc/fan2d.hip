// Fan beam with few slices (DD2_multiGPU, src/multiSlices_ker.cuh:35-38,
// methods 0 and 2, with up to kFanSlices slices; slice v of the detector sees
// slice v of the image only): the branchless DD projector and backprojector
// with lanes over detector boundaries and over row boundaries.
//
// The general walks (dd_device.hpp) hold lane = z, so with one slice they
// use one lane in 64 (16 with 16 slices).  Here the same quantities are formed
// with the lanes spread the other way, every slice in turn from the same
// geometry (row coordinates, columns, fractions):
//   projector: the row prefix C of every slab (C(m) = sum of the slab's first
//     m voxels, piecewise linear in between) is built once per call, in both
//     slab orientations.  Column u's footprint integral on slab j is
//     C(a_{u+1}) - C(a_u) at its two boundaries' row coordinates a (clamped
//     to the slab), which is what the walk's running row sums give; a block
//     holds 64 boundaries of one view (63 columns), each lane reads C at its
//     own boundary and takes its upper neighbour's by a DPP shift.
//   backprojector: the walk's running column sum W at a row boundary is an
//     interpolation of the columns' prefix sum Q at the column that holds
//     the boundary (from a one-row start table), so a wavefront holds 64 row
//     boundaries of one slab, reads per view only the columns its boundaries
//     fall in, and differences W with its upper neighbour by a DPP shift.
// Factors, clamping and the slab sums are those of the general walks, so the
// two agree to rounding.
#include "dd_device.hpp"
#include "plan.hpp"

#include <algorithm>

namespace ctlab {

bool is_fan2d(const DevGeom& g) { return !g.cone && g.ZN == g.DNV && g.ZN <= kFanSlices; }

// Slice z's row prefixes start at z * fan_prefix_elems(g) in the integral-image
// workspace: Py (YN rows of XN + 1 entries), then Px (XN of YN + 1).
__host__ __device__ inline long fan_prefix_elems(const DevGeom& g)
{
    return (long)g.YN * (g.XN + 1) + (long)g.XN * (g.YN + 1);
}

// Masked row prefixes of every slice z for both orientations: Py[y][m] =
// sum_{x<m} f(x, y, z) (XN + 1 entries per y), Px[x][m] = sum_{y<m} f(x, y, z)
// (YN + 1 entries per x).  One wavefront per (slice, row or column), 64
// entries per pass with the running total carried in a register.
template <typename T>
__global__ void __launch_bounds__(256) fan2d_prefix_kernel(DevGeom g, const float* __restrict__ vol, T* __restrict__ C)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int XN = g.XN, YN = g.YN, ZN = g.ZN;
    if (w >= (XN + YN) * ZN) return;
    const int z = w / (XN + YN), wr = w - z * (XN + YN);
    T* Py = C + z * fan_prefix_elems(g);
    T* Px = Py + (long)YN * (XN + 1);
    const bool yrow = wr < YN;
    const int j = yrow ? wr : wr - YN;
    const int n = yrow ? XN : YN;
    T* out = yrow ? Py + (long)j * (XN + 1) : Px + (long)j * (YN + 1);
    if (lane == 0) out[0] = T(0);
    T carry = T(0);
    for (int m0 = 0; m0 < n; m0 += 64) {
        const int m = m0 + lane;
        T v = T(0);
        if (m < n) {
            const long col = yrow ? (long)j * XN + m : (long)m * XN + j;
            const T mk = (g.mask == nullptr || g.mask[col]) ? T(1) : T(0);
            v = (T)vol[col * ZN + z] * mk;
        }
        const T r = wave_inclusive_scan(v) + carry;
        if (m < n) out[m + 1] = r;
        carry = readlane_t(r, 63);
    }
}

// Projector: block (view p, tile of 63 columns from u0), lane l holds
// boundary u0 + l; the block's four wavefronts take every fourth slab (a
// wavefront's reads are a chain of slabs, so four chains finish sooner, and a
// 36-view subset still fills the GPU), and their sums are added in LDS in a
// fixed order.  SL slices (SL <= SLMAX, a compile-time bound) share each
// slab's row coordinates.  The output factor is the general epilogue's.
// Above kFanProjSlices slices (two reads per slice and slab in each lane) the
// lane = z projector, with 9-16 of its lanes busy, is as fast (measured at
// 256x256, 360 views: 1.81 against 1.75 ms at 16 slices, 0.26 against 1.79 ms
// at 4), so the projector takes this path up to kFanProjSlices slices.
constexpr int kFanCols = 63;

template <typename T, int SLMAX>
__global__ void __launch_bounds__(256) fan2d_proj_kernel(DevGeom g, const T* __restrict__ Cz,
                                                         float* __restrict__ prj, int ntiles)
{
    __shared__ double sacc[4][SLMAX][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x / ntiles, tile = blockIdx.x - p * ntiles;
    const int DNU = g.DNU, SL = g.ZN;
    const int u0 = tile * kFanCols, nu = min(kFanCols, DNU - u0);
    const ViewParams vp = g.vp[p];
    const double kap = g.kap[(long)p * (DNU + 1) + u0 + min(lane, nu)];
    const int NR = vp.NR, NS = vp.NS;
    const long pre = fan_prefix_elems(g);
    const T* __restrict__ C = Cz + (vp.yslab ? 0L : (long)g.YN * (g.XN + 1));   // slice 0's prefixes
    const double* __restrict__ invs = g.invs + (long)p * g.sjs;
    // the boundary's row coordinate at slab wv, then four slabs per step
    double a = vp.rbase + (vp.s0 * (1.0 / g.dx) + wv) * kap;
    const double kap4 = 4.0 * kap;
    double acc[SLMAX];
#pragma unroll
    for (int z = 0; z < SLMAX; ++z) acc[z] = 0.0;
#pragma unroll 4
    for (int j = wv; j < NS; j += 4) {
        const double al = fmin(fmax(a, 0.0), (double)NR);
        const int m = min((int)al, NR - 1);
        const T fr = (T)(al - (double)m);
        const double qj = fabs(invs[j]);
#pragma unroll
        for (int z = 0; z < SLMAX; ++z) {
            if (SLMAX == 1 || z < SL) {
                const T* __restrict__ row = C + z * pre + (long)j * (NR + 1);
                const T c0 = row[m], c1 = row[m + 1];
                const T Cb = c0 + fr * (c1 - c0);
                const T Cn = shift_down1(Cb);                 // boundary l + 1 (lane 63's is not used)
                const T H = vp.asc ? Cn - Cb : Cb - Cn;       // footprint integral of column l
                acc[z] += (double)H * qj;
            }
        }
        a += kap4;
    }
#pragma unroll
    for (int z = 0; z < SLMAX; ++z) sacc[wv][z][lane] = acc[z];
    __syncthreads();
    if (wv != 0 || lane >= nu) return;
    const long cu = (long)p * DNU + u0 + lane;
    const double cf = g.colfac[cu], e2 = g.exy2[cu];
    for (int z = 0; z < SL; ++z) {
        const double t = ((sacc[0][z][lane] + sacc[1][z][lane]) + sacc[2][z][lane]) + sacc[3][z][lane];
        const double ez = g.ez[z];
        prj[cu * g.DNV + z] = (float)(t * cf * sqrt(e2 + ez * ez) * g.rowfac[z]);
    }
}

// Backprojector.  For a view and slab the walk's running column sum W at row
// boundary r (row-axis coordinate) is, with t the walk-order column that
// holds r, W(r) = Q[t] + fr * (Q[t + 1] - Q[t]): Q[c] is the sum of the
// column entries (y * L with one detector row) of walk-order columns below c,
// and fr = clamp((km - klo_t) / |slope step_t|, 0, 1) with km = (r - rbase) /
// |rs| in slope space, the walk's own fraction.  Voxel row m of the slab gets
// W(m + 1) - W(m).  So every lane of a wavefront (slab, tile of 63 rows) holds
// one row boundary and, per view, reads only the column that holds it (from a
// start table with one entry per row boundary, built once per plan) and two
// prefix entries, instead of walking every column the rows cross; its
// upper neighbour's W comes by a DPP shift.  Block = one (slab, tile); its
// four wavefronts take every fourth view of the list, and their sums are added
// in LDS in a fixed order.  The first nby blocks take the y-slab views and
// write vol_y, the rest the x-slab views into vol_x (summed by fan2d_add).
constexpr int kFanRows = 63;

// start1[(p*SJ + j)*SR1 + r]: the walk-order column whose extent holds row
// boundary r of slab j (the first column whose upper edge lies above r), for
// r = 0 .. NR; 0xFFFF where every column ends at or below r (back_start_kernel
// at one-row resolution, including the slab's far edge r = NR).
__global__ void __launch_bounds__(256) fan2d_start_kernel(DevGeom g, const ColRec* __restrict__ rec,
                                                          unsigned short* __restrict__ start1, int SJ, int SR1)
{
    const long idx = blockIdx.x * 256L + threadIdx.x;
    if (idx >= (long)g.PN * SJ * SR1) return;
    const int r = (int)(idx % SR1);
    const long pj = idx / SR1;
    const int j = (int)(pj % SJ);
    const int p = (int)(pj / SJ);
    const ViewParams vp = g.vp[p];
    const int DNU = g.DNU;
    if (j >= vp.NS || r > vp.NR) { start1[idx] = 0xFFFF; return; }
    const double rs = fabs((vp.s0 + j * g.dx) / g.dx);
    const ColRec* R = rec + (long)p * (DNU + 1);
    auto A = [&](int q) { return vp.rbase + rs * (q < DNU ? R[q].klo : R[DNU - 1].khi); };
    const double rr = (double)r;
    if (A(DNU) <= rr) { start1[idx] = 0xFFFF; return; }
    int lo = 1, hi = DNU;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A(mid) > rr) hi = mid; else lo = mid + 1;
    }
    start1[idx] = (unsigned short)(lo - 1);
}

// Q[(z*PN + p)*(DNU + 1) + c] = sum of y * L over walk-order columns below c
// (c = 0 .. DNU) of detector row (slice) z, L as prjprefix_kernel's; one
// wavefront per (slice, view), 64 columns per pass
template <typename T>
__global__ void __launch_bounds__(256) fan2d_colprefix_kernel(DevGeom g, const float* __restrict__ prj,
                                                              T* __restrict__ Q)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= g.PN * g.ZN) return;
    const int z = w / g.PN, p = w - z * g.PN;
    const int DNU = g.DNU, asc = g.vp[p].asc;
    const double ez = g.ez[z];
    T* Qv = Q + (long)w * (DNU + 1);
    if (lane == 0) Qv[0] = T(0);
    T carry = T(0);
    for (int c0 = 0; c0 < DNU; c0 += 64) {
        const int c = c0 + lane;
        T v = T(0);
        if (c < DNU) {
            const long cu = (long)p * DNU + (asc ? c : DNU - 1 - c);
            const double L = g.dx * sqrt(g.exy2[cu] + ez * ez) * fabs(g.tau[cu]);
            v = (T)((double)prj[cu * g.DNV + z] * L);
        }
        const T r = wave_inclusive_scan(v) + carry;
        if (c < DNU) Qv[c + 1] = r;
        carry = readlane_t(r, 63);
    }
}

template <typename T, int SLMAX>
__global__ void __launch_bounds__(256) fan2d_back_kernel(DevGeom g, const T* __restrict__ Q, float* __restrict__ vol_y,
                                                         float* __restrict__ vol_x, const int* __restrict__ views_y,
                                                         int ny, const int* __restrict__ views_x, int nx, int nby,
                                                         int ntr_y, int ntr_x,
                                                         const unsigned short* __restrict__ start1, int SJ, int SR1)
{
    __shared__ double sacc[4][SLMAX][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ys = blockIdx.x < (unsigned)nby;
    const int b = ys ? blockIdx.x : blockIdx.x - nby;
    const int NR = ys ? g.XN : g.YN, ntr = ys ? ntr_y : ntr_x;
    const int rt = b % ntr;
    const int j = b / ntr;
    const int* __restrict__ views = ys ? views_y : views_x;
    const int nviews = ys ? ny : nx;
    float* __restrict__ vol = ys ? vol_y : vol_x;
    const int r0 = rt * kFanRows;
    const int nm = min(kFanRows, NR - r0);
    const int r = min(r0 + lane, NR);            // this lane's row boundary (lanes past the tile repeat the last)
    const int DNU = g.DNU, SL = g.ZN;
    const long qslice = (long)g.PN * (DNU + 1);
    double acc[SLMAX];
#pragma unroll
    for (int z = 0; z < SLMAX; ++z) acc[z] = 0.0;
#pragma unroll 2
    for (int iv = wv; iv < nviews; iv += 4) {
        const int p = __builtin_amdgcn_readfirstlane(views[iv]);
        const ViewParams vp = g.vp[p];
        const double inv_rs = fabs(g.dx * g.invs[(long)p * g.sjs + j]);
        const double km = ((double)r - vp.rbase) * inv_rs;
        const int t0 = start1[((long)p * SJ + j) * SR1 + r];
        const int t = t0 == 0xFFFF ? DNU : t0;
        // column t's lower slope (times sign(s0)) and 1/|slope step|, the
        // record's values, read from the natural-order tables (8 bytes per
        // column: neighbouring lanes' columns are a few entries apart); past
        // the last column (t = DNU, every column below r) the fraction is 0
        const double sgn_s = vp.s0 > 0.0 ? 1.0 : -1.0;
        const int asc = vp.asc;
        const double klo = sgn_s * g.kap[(long)p * (DNU + 1) + (asc ? t : DNU - t)];
        const double iw = t < DNU ? g.invdk[(long)p * DNU + (asc ? t : DNU - 1 - t)] : 0.0;
        const double fr = fmin(fmax((km - klo) * iw, 0.0), 1.0);
        const int t1 = min(t + 1, DNU);
#pragma unroll
        for (int z = 0; z < SLMAX; ++z) {
            if (SLMAX == 1 || z < SL) {
                const T* __restrict__ Qv = Q + z * qslice + (long)p * (DNU + 1);
                const T q0 = Qv[t], q1 = Qv[t1];
                const double W = (double)q0 + fr * (double)(q1 - q0);
                acc[z] += shift_down1(W) - W;    // lane l: W(r + 1) - W(r)
            }
        }
    }
#pragma unroll
    for (int z = 0; z < SLMAX; ++z) sacc[wv][z][lane] = acc[z];
    __syncthreads();
    if (wv != 0 || lane >= nm) return;
    const double sgn = g.dzh < 0.0 ? -1.0 : 1.0;
    const int m = r0 + lane;
    const long col = ys ? ((long)j * g.XN + m) : ((long)m * g.XN + j);
    const float mk = (g.mask == nullptr || g.mask[col]) ? 1.0f : 0.0f;
    for (int z = 0; z < SL; ++z) {
        const double t = ((sacc[0][z][lane] + sacc[1][z][lane]) + sacc[2][z][lane]) + sacc[3][z][lane];
        vol[col * g.ZN + z] = (float)(t * sgn) * mk;
    }
}

// vol += add (n floats)
__global__ void __launch_bounds__(256) fan2d_add(float* __restrict__ vol, const float* __restrict__ add, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        vol[i] += add[i];
}

template <typename T>
hipError_t launch_fan2d_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s)
{
    T* C = static_cast<T*>(d_cz);
    const long nw = (long)(g.XN + g.YN) * g.ZN;
    hipLaunchKernelGGL(fan2d_prefix_kernel<T>, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, s, g, d_vol, C);
    const int ntiles = (g.DNU + kFanCols - 1) / kFanCols;
    const dim3 grid((unsigned)g.PN * (unsigned)ntiles);
    if (g.ZN == 1) hipLaunchKernelGGL((fan2d_proj_kernel<T, 1>), grid, dim3(256), 0, s, g, C, d_prj, ntiles);
    else if (g.ZN <= 4) hipLaunchKernelGGL((fan2d_proj_kernel<T, 4>), grid, dim3(256), 0, s, g, C, d_prj, ntiles);
    else hipLaunchKernelGGL((fan2d_proj_kernel<T, kFanProjSlices>), grid, dim3(256), 0, s, g, C, d_prj, ntiles);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fan2d_back(const DevGeom& g, const float* d_prj, void* d_Q, float* d_vol, float* d_scratch,
                             const int* d_views_y, int ny, const int* d_views_x, int nx,
                             const unsigned short* d_start1, int SJ, int SR1, hipStream_t s)
{
    T* Q = static_cast<T*>(d_Q);
    const long nq = (long)g.PN * g.ZN;
    hipLaunchKernelGGL(fan2d_colprefix_kernel<T>, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, g, d_prj, Q);
    const int ntr_y = (g.XN + kFanRows - 1) / kFanRows, ntr_x = (g.YN + kFanRows - 1) / kFanRows;
    const int nby = ntr_y * g.YN, nbx = ntr_x * g.XN;
    const dim3 grid((unsigned)(nby + nbx));
    if (g.ZN == 1)
        hipLaunchKernelGGL((fan2d_back_kernel<T, 1>), grid, dim3(256), 0, s, g, Q, d_vol, d_scratch, d_views_y, ny,
                           d_views_x, nx, nby, ntr_y, ntr_x, d_start1, SJ, SR1);
    else if (g.ZN <= 4)
        hipLaunchKernelGGL((fan2d_back_kernel<T, 4>), grid, dim3(256), 0, s, g, Q, d_vol, d_scratch, d_views_y, ny,
                           d_views_x, nx, nby, ntr_y, ntr_x, d_start1, SJ, SR1);
    else
        hipLaunchKernelGGL((fan2d_back_kernel<T, kFanSlices>), grid, dim3(256), 0, s, g, Q, d_vol, d_scratch,
                           d_views_y, ny, d_views_x, nx, nby, ntr_y, ntr_x, d_start1, SJ, SR1);
    const long nvox = (long)g.XN * g.YN * g.ZN;
    hipLaunchKernelGGL(fan2d_add, dim3((unsigned)std::min<long>((nvox + 255) / 256, 8192)), dim3(256), 0, s, d_vol,
                       d_scratch, nvox);
    return hipGetLastError();
}

hipError_t launch_fan2d_start(const DevGeom& g, const void* d_rec, unsigned short* d_start1, int SJ, int SR1,
                              hipStream_t s)
{
    const long n = (long)g.PN * SJ * SR1;
    hipLaunchKernelGGL(fan2d_start_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g,
                       (const ColRec*)d_rec, d_start1, SJ, SR1);
    return hipGetLastError();
}

template hipError_t launch_fan2d_proj<double>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_fan2d_proj<float>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_fan2d_back<double>(const DevGeom&, const float*, void*, float*, float*, const int*, int,
                                              const int*, int, const unsigned short*, int, int, hipStream_t);
template hipError_t launch_fan2d_back<float>(const DevGeom&, const float*, void*, float*, float*, const int*, int,
                                             const int*, int, const unsigned short*, int, int, hipStream_t);

}  // namespace ctlab

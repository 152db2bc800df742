// Single-slice fan beam (DD2_multiGPU with one slice, src/multiSlices_ker.cuh:
// 35-38, methods 0 and 2): the branchless DD projector and backprojector with
// lanes over detector boundaries and over views.
//
// The general walks (dd_device.hpp) hold lane = z, so with one slice they
// use one lane in 64.  Here the same quantities are formed with the lanes
// spread the other way:
//   projector: the row prefix C of every slab (C(m) = sum of the slab's first
//     m voxels, piecewise linear in between) is built once per call, in both
//     slab orientations.  Column u's footprint integral on slab j is
//     C(a_{u+1}) - C(a_u) at its two boundaries' row coordinates a (clamped
//     to the slab), which is what the walk's running row sums give; a block
//     holds 64 boundaries of one view (63 columns), each lane reads C at its
//     own boundary and takes its upper neighbour's by a DPP shift.
//   backprojector: the walk of back_walk_accum (column records in slope
//     space, start table, running column sums W at the row boundaries) with
//     lane = view: a wavefront (slab, tile of NM rows) walks 64 views of one
//     slab orientation at once, each lane its own view, and the rows' sums
//     over the views are reduced across lanes at the end in a fixed order.
// Factors, clamping and the slab sums are those of the general walks, so the
// two agree to rounding.
#include "dd_device.hpp"
#include "plan.hpp"

#include <algorithm>

namespace ctlab {

bool is_fan2d(const DevGeom& g) { return !g.cone && g.ZN == 1 && g.DNV == 1; }

// Masked row prefixes of the image for both orientations: Py[y][m] =
// sum_{x<m} f(x, y) (XN + 1 entries per y), Px[x][m] = sum_{y<m} f(x, y)
// (YN + 1 entries per x).  One wavefront per row or column, 64 entries per
// pass with the running total carried in a register.
template <typename T>
__global__ void __launch_bounds__(256) fan2d_prefix_kernel(DevGeom g, const float* __restrict__ vol,
                                                           T* __restrict__ Py, T* __restrict__ Px)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int XN = g.XN, YN = g.YN;
    if (w >= XN + YN) return;
    const bool yrow = w < YN;
    const int j = yrow ? w : w - YN;
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
            v = (T)vol[col] * mk;
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
// fixed order.  The output factor is the general epilogue's (one row).
constexpr int kFanCols = 63;

template <typename T>
__global__ void __launch_bounds__(256) fan2d_proj_kernel(DevGeom g, const T* __restrict__ Py,
                                                         const T* __restrict__ Px, float* __restrict__ prj,
                                                         int ntiles)
{
    __shared__ double sacc[4][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x / ntiles, tile = blockIdx.x - p * ntiles;
    const int DNU = g.DNU;
    const int u0 = tile * kFanCols, nu = min(kFanCols, DNU - u0);
    const ViewParams vp = g.vp[p];
    const double kap = g.kap[(long)p * (DNU + 1) + u0 + min(lane, nu)];
    const int NR = vp.NR, NS = vp.NS;
    const T* __restrict__ C = vp.yslab ? Py : Px;
    const double* __restrict__ invs = g.invs + (long)p * g.sjs;
    // the boundary's row coordinate at slab wv, then four slabs per step
    double a = vp.rbase + (vp.s0 * (1.0 / g.dx) + wv) * kap;
    const double kap4 = 4.0 * kap;
    double acc = 0.0;
#pragma unroll 4
    for (int j = wv; j < NS; j += 4) {
        const double al = fmin(fmax(a, 0.0), (double)NR);
        const int m = min((int)al, NR - 1);
        const T fr = (T)(al - (double)m);
        const T* __restrict__ row = C + (long)j * (NR + 1);
        const T c0 = row[m], c1 = row[m + 1];
        const T Cb = c0 + fr * (c1 - c0);
        const T Cn = shift_down1(Cb);                 // boundary l + 1 (lane 63's is not used)
        const T H = vp.asc ? Cn - Cb : Cb - Cn;       // footprint integral of column l
        acc += (double)H * fabs(invs[j]);
        a += kap4;
    }
    sacc[wv][lane] = acc;
    __syncthreads();
    if (wv != 0) return;
    acc = ((sacc[0][lane] + sacc[1][lane]) + sacc[2][lane]) + sacc[3][lane];
    if (lane < nu) {
        const long cu = (long)p * DNU + u0 + lane;
        const double ez = g.ez[0], rf = g.rowfac[0];
        prj[cu] = (float)(acc * g.colfac[cu] * sqrt(g.exy2[cu] + ez * ez) * rf);
    }
}

// Backprojector: block = 4 wavefronts = 4 consecutive slabs of one tile of
// kFanRows rows; the first nby blocks take the y-slab views and write vol_y,
// the rest the x-slab views into vol_x (summed by back_add), as dd_back_walk.
constexpr int kFanRows = 16;

template <typename T>
__global__ void __launch_bounds__(256) fan2d_back_kernel(DevGeom g, const T* __restrict__ P, float* __restrict__ vol_y,
                                                         float* __restrict__ vol_x, const int* __restrict__ views_y,
                                                         int ny, const int* __restrict__ views_x, int nx, int nby,
                                                         int ntr_y, int ntr_x, const ColRec* __restrict__ rec,
                                                         const unsigned short* __restrict__ start, int SJ, int SR)
{
    constexpr int NM = kFanRows;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ys = blockIdx.x < (unsigned)nby;
    const int b = ys ? blockIdx.x : blockIdx.x - nby;
    const int NS = ys ? g.YN : g.XN, NR = ys ? g.XN : g.YN, ntr = ys ? ntr_y : ntr_x;
    const int rt = b % ntr;
    const int j = (b / ntr) * 4 + wv;
    if (j >= NS) return;
    const int* __restrict__ views = ys ? views_y : views_x;
    const int nviews = ys ? ny : nx;
    float* __restrict__ vol = ys ? vol_y : vol_x;
    const int r0 = rt * NM;
    const int nm = min(NM, NR - r0);
    const int DNU = g.DNU;
    const long srow = (long)j * SR + (r0 >> 3);
    const long sstride = (long)SJ * SR;
    double acc[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] = 0.0;

    for (int base = 0; base < nviews; base += 64) {
        const int iv = base + lane;
        const int p = iv < nviews ? views[iv] : views[0];
        const int st = iv < nviews ? (int)start[p * sstride + srow] : 0xFFFF;
        bool act = st != 0xFFFF;
        const int stc = act ? st : 0;
        const ViewParams vp = g.vp[p];
        const double inv_sj = g.invs[(long)p * g.sjs + j];
        const double inv_rs = fabs(g.dx * inv_sj);
        // row boundary r0 + i in slope space, km = (r0 + i - rbase) / |rs|
        double km = ((double)r0 - vp.rbase) * inv_rs;
        const ColRec* __restrict__ rq = rec + (long)p * (DNU + 1) + stc;
        const ColRec r = rq[0];
        act = act && !(r.klo >= fma((double)nm, inv_rs, km));
        // a view's v-prefix entries in walk order (one row: the entry is y * L itself);
        // the sentinel column past either end reads 0
        const T* __restrict__ Pv = P + (long)p * DNU;
        const int asc = vp.asc;
        auto pcol = [&](int c) { return c < DNU ? Pv[asc ? c : DNU - 1 - c] : T(0); };
        double klo = r.klo, khi = r.khi, iw = r.invdk;
        int c = stc;
        T D = pcol(c);
        T Dn = pcol(c + 1);
        double rn_khi = rq[1].khi, rn_iw = rq[1].invdk;
        T Tsum = T(0), Wprev = T(0);
#pragma unroll
        for (int i = 0; i <= NM; ++i) {
            if (i <= nm) {
                if (act) {
                    // advance while the column ends at or below this row boundary;
                    // the sentinel's NaN upper edge stops the walk
                    while (km >= khi) {
                        Tsum += D;
                        klo = khi;
                        khi = rn_khi;
                        iw = rn_iw;
                        D = Dn;
                        ++c;
                        rn_khi = rq[c + 1 - stc].khi;
                        rn_iw = rq[c + 1 - stc].invdk;
                        Dn = pcol(c + 1);
                    }
                    const double fr = fmin(fmax((km - klo) * iw, 0.0), 1.0);
                    km += inv_rs;
                    const T W = Tsum + (T)fr * D;
                    if (i > 0) acc[i - 1] += (double)(W - Wprev);
                    Wprev = W;
                }
            }
        }
    }
    // rows' sums over the views (lane order), then lane i stores row i
    const double sgn = g.dzh < 0.0 ? -1.0 : 1.0;
    double mine = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const double tot = readlane_d(wave_inclusive_scan(acc[i]), 63);
        if (lane == i) mine = tot;
    }
    if (lane < nm) {
        const int m = r0 + lane;
        const long col = ys ? ((long)j * g.XN + m) : ((long)m * g.XN + j);
        const float mk = (g.mask == nullptr || g.mask[col]) ? 1.0f : 0.0f;
        vol[col] = (float)(mine * sgn) * mk;
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
    T* Py = static_cast<T*>(d_cz);
    T* Px = Py + (long)g.YN * (g.XN + 1);
    hipLaunchKernelGGL(fan2d_prefix_kernel<T>, dim3((unsigned)((g.XN + g.YN + 3) / 4)), dim3(256), 0, s, g, d_vol,
                       Py, Px);
    const int ntiles = (g.DNU + kFanCols - 1) / kFanCols;
    hipLaunchKernelGGL(fan2d_proj_kernel<T>, dim3((unsigned)g.PN * (unsigned)ntiles), dim3(256), 0, s, g, Py, Px,
                       d_prj, ntiles);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fan2d_back(const DevGeom& g, const T* P, float* d_vol, float* d_scratch, const int* d_views_y,
                             int ny, const int* d_views_x, int nx, const void* d_rec, const unsigned short* d_start,
                             int SJ, int SR, hipStream_t s)
{
    const int ntr_y = (g.XN + kFanRows - 1) / kFanRows, ntr_x = (g.YN + kFanRows - 1) / kFanRows;
    const int nby = ntr_y * ((g.YN + 3) / 4), nbx = ntr_x * ((g.XN + 3) / 4);
    hipLaunchKernelGGL(fan2d_back_kernel<T>, dim3((unsigned)(nby + nbx)), dim3(256), 0, s, g, P, d_vol, d_scratch,
                       d_views_y, ny, d_views_x, nx, nby, ntr_y, ntr_x, (const ColRec*)d_rec, d_start, SJ, SR);
    const long nvox = (long)g.XN * g.YN;
    hipLaunchKernelGGL(fan2d_add, dim3((unsigned)std::min<long>((nvox + 255) / 256, 8192)), dim3(256), 0, s, d_vol,
                       d_scratch, nvox);
    return hipGetLastError();
}

template hipError_t launch_fan2d_proj<double>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_fan2d_proj<float>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_fan2d_back<double>(const DevGeom&, const double*, float*, float*, const int*, int,
                                              const int*, int, const void*, const unsigned short*, int, int,
                                              hipStream_t);
template hipError_t launch_fan2d_back<float>(const DevGeom&, const float*, float*, float*, const int*, int,
                                             const int*, int, const void*, const unsigned short*, int, int,
                                             hipStream_t);

}  // namespace ctlab

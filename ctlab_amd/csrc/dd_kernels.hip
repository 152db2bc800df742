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
#include "dd_common.hpp"

namespace ctlab {

// Intra-wavefront hand-off through LDS: every lane writes, then every lane
// reads other lanes' values.  LDS executes one wavefront's instructions in
// order; the fences keep the compiler from moving the accesses.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// piecewise-linear read of a prefix table G[0..n] (G[0] = 0) at b in [0, n]
template <typename T>
__device__ __forceinline__ T lds_interp(const T* G, double b, int n)
{
    b = fmin(fmax(b, 0.0), (double)n);
    const int i = min((int)b, n - 1);
    const T fr = (T)(b - (double)i);
    const T g0 = G[i];
    const T g1 = G[i + 1];
    return g0 + fr * (g1 - g0);
}

// ---------------------------------------------------------------------------
// z-prefix of every masked voxel column: cz[col*ZN + k] = sum_{k'<=k} f*mask
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
zprefix_kernel(const float* __restrict__ vol, T* __restrict__ cz,
               const unsigned char* __restrict__ mask, long ncol, int ZN)
{
    const long col = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const T mk = (mask == nullptr || mask[col]) ? T(1) : T(0);
    const float* __restrict__ src = vol + col * ZN;
    T* __restrict__ dst = cz + col * ZN;
    T run = T(0);
    for (int k = 0; k < ZN; ++k) {
        run += (T)src[k] * mk;
        dst[k] = run;
    }
}

// ---------------------------------------------------------------------------
// Projector walk.  Block = 4 wavefronts = 4 consecutive views of one tile of
// NU detector columns (neighbouring views read nearly the same slab rows);
// each wavefront loops over all slabs of its view.  Lane = z for the SAT rows,
// lane = v for outputs.
// ---------------------------------------------------------------------------
template <typename T, int NU>
__global__ void __launch_bounds__(256)
dd_proj_walk(DevGeom g, const T* __restrict__ cz, float* __restrict__ prj)
{
    __shared__ T sG[4][72];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = blockIdx.y * 4 + wv;
    if (p >= g.PN) return;
    T* G = sG[wv];
    if (lane == 0) G[0] = T(0);
    const ViewParams vp = g.vp[p];
    const double* __restrict__ kap = g.kap + (long)p * (g.DNU + 1);
    const double* __restrict__ tau = g.tau + (long)p * g.DNU;
    const int u0 = blockIdx.x * NU;
    const int nu = min(NU, g.DNU - u0);
    const int ZN = g.ZN, NR = vp.NR, DNV = g.DNV;
    const int zl = min(lane, ZN - 1);
    const int vl = min(lane, DNV - 1);
    const double zlo = g.zhat[vl], zhi = g.zhat[vl + 1];
    const double inv_dx = 1.0 / g.dx;
    const long slab_off = (long)vp.slab_stride * ZN;
    const long row_off = (long)vp.row_stride * ZN;
    const long last = (long)(NR - 1) * row_off;

    T acc[NU];
#pragma unroll
    for (int q = 0; q < NU; ++q) acc[q] = T(0);

    for (int j = 0; j < vp.NS; ++j) {
        const double sj = vp.s0 + j * g.dx;
        const double rs = sj * inv_dx;
        const double afirst = vp.rbase + rs * kap[vp.asc ? u0 : u0 + nu];
        const double alast = vp.rbase + rs * kap[vp.asc ? u0 + nu : u0];
        if (alast <= 0.0 || afirst >= (double)NR) continue;
        const T qj = (T)(g.cone ? 1.0 / (sj * sj) : 1.0 / fabs(sj));
        const T* __restrict__ base = cz + j * slab_off + zl;
        double ap = fmin(fmax(afirst, 0.0), (double)NR);
        int m = min((int)ap, NR - 1);
        double mf = (double)m;
        long off = m * row_off;
        // rows m, m+1, m+2 in flight (indices clamped to the last row)
        T c0 = base[off];
        T c1 = base[min(off + row_off, last)];
        T c2 = base[min(off + 2 * row_off, last)];
        T S = T(0);
        T Eprev = (T)(ap - mf) * c0;
        for (int q = 1; q <= nu; ++q) {
            double aq = vp.rbase + rs * kap[vp.asc ? u0 + q : u0 + nu - q];
            aq = fmin(fmax(aq, 0.0), (double)NR);
            while (m < NR - 1 && aq >= mf + 1.0) {
                S += c0;
                ++m;
                mf += 1.0;
                off += row_off;
                c0 = c1;
                c1 = c2;
                c2 = base[min(off + 2 * row_off, last)];
            }
            const T E = S + (T)(aq - mf) * c0;
            const T H = E - Eprev;
            Eprev = E;
            if (aq > ap) {
                const int u = vp.asc ? u0 + q - 1 : u0 + nu - q;
                const double t = g.cone ? sj * tau[u] : 1.0;
                G[lane + 1] = H;
                wave_lds_sync();
                const T Dhi = lds_interp(G, vp.beta0 + t * zhi, ZN);
                const T Dlo = lds_interp(G, vp.beta0 + t * zlo, ZN);
                acc[q - 1] += (Dhi - Dlo) * qj;
                wave_lds_sync();
            }
            ap = aq;
        }
    }

    if (lane < DNV) {
        const double ezv = g.ez[lane];
        const double rf = g.rowfac[lane];
        float* __restrict__ out = prj + (long)p * g.DNU * DNV + lane;
#pragma unroll
        for (int q = 0; q < NU; ++q) {
            if (q >= nu) break;
            const int u = vp.asc ? u0 + q : u0 + nu - 1 - q;
            const long cu = (long)p * g.DNU + u;
            const double val = (double)acc[q] * g.colfac[cu] * sqrt(g.exy2[cu] + ezv * ezv) * rf;
            out[(long)u * DNV] = (float)val;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-(view, column) inclusive v-prefix of y*L for the backprojector.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
prjprefix_kernel(DevGeom g, const float* __restrict__ prj, T* __restrict__ P)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;  // p*DNU + u
    if (idx >= (long)g.PN * g.DNU) return;
    const double ta = fabs(g.tau[idx]);
    const double e2 = g.exy2[idx];
    const float* __restrict__ src = prj + idx * g.DNV;
    T* __restrict__ dst = P + idx * g.DNV;
    T run = T(0);
    for (int v = 0; v < g.DNV; ++v) {
        const double ezv = g.ez[v];
        const double L = g.dx * sqrt(e2 + ezv * ezv) * ta;
        run += (T)((double)src[v] * L);
        dst[v] = run;
    }
}

// Adjoint z-part for detector column u: sum_v y*L * (fraction of cell v covered
// by the voxel z-interval [lane, lane+1] mapped onto the detector rows), read
// from the column's v-prefix Pu (lane v holds prefix v+1) through LDS.
template <typename T>
__device__ __forceinline__ T column_delta(const DevGeom& g, T* G, T Pu, int p, int u, double inv_sj,
                                          double beta0, int lane)
{
    const int DNV = g.DNV;
    G[lane + 1] = Pu;
    wave_lds_sync();
    double nlo, nhi;
    T sgn = T(1);
    if (g.cone) {
        const double inv = inv_sj * g.esdz[(long)p * g.DNU + u];   // 1 / (t * dzh)
        const double off = -beta0 * inv - g.zhat0 / g.dzh;
        nlo = (double)lane * inv + off;
        nhi = nlo + inv;
        if (g.dzh < 0.0) sgn = T(-1);
    } else {
        nlo = (double)lane;
        nhi = (double)(lane + 1);
    }
    const T r = (lds_interp(G, nhi, DNV) - lds_interp(G, nlo, DNV)) * sgn;
    wave_lds_sync();
    return r;
}

// Block = 4 wavefronts = 4 consecutive slabs of one tile of NM rows (their
// detector-column ranges nearly coincide); each wavefront loops over the views
// of one slab orientation; lane = z.
template <typename T, int NM>
__global__ void __launch_bounds__(256)
dd_back_walk(DevGeom g, const T* __restrict__ P, float* __restrict__ vol,
             const int* __restrict__ views, int nviews, int yslab, int accumulate,
             int NS, int NR, int ntr)
{
    __shared__ T sG[4][72];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rt = blockIdx.x % ntr;
    const int j = (blockIdx.x / ntr) * 4 + wv;
    if (j >= NS) return;
    T* G = sG[wv];
    if (lane == 0) G[0] = T(0);
    const int r0 = rt * NM;
    const int nm = min(NM, NR - r0);
    const int ZN = g.ZN, DNU = g.DNU, DNV = g.DNV;
    const int vl = min(lane, DNV - 1);
    const double inv_dx = 1.0 / g.dx;

    T acc[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] = T(0);

    for (int iv = 0; iv < nviews; ++iv) {
        const int p = views[iv];
        const ViewParams vp = g.vp[p];
        const double* __restrict__ kap = g.kap + (long)p * (DNU + 1);
        const double* __restrict__ invdk = g.invdk + (long)p * DNU;
        const T* __restrict__ Pv = P + (long)p * DNU * DNV + vl;
        const double sj = vp.s0 + j * g.dx;
        const double rs = sj * inv_dx;
        const int asc = vp.asc;
        auto A = [&](int q) { return vp.rbase + rs * kap[asc ? q : DNU - q]; };
        const double A0 = A(0), AN = A(DNU);
        if (AN <= (double)r0 || A0 >= (double)(r0 + nm)) continue;
        int lo = 1, hi = DNU;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A(mid) > (double)r0) hi = mid; else lo = mid + 1;
        }
        int c = lo - 1;
        const double inv_sj = 1.0 / sj;
        const double inv_rs = fabs(g.dx * inv_sj);
        auto ucol = [&](int cc) { return asc ? cc : DNU - 1 - cc; };
        // P rows of columns c and c+1 in flight
        T Pr0 = Pv[(long)ucol(c) * DNV];
        T Pr1 = Pv[(long)ucol(min(c + 1, DNU - 1)) * DNV];
        double alo = A(c), ahi = A(c + 1);
        double iw = inv_rs * invdk[ucol(c)];
        T D = column_delta(g, G, Pr0, p, ucol(c), inv_sj, vp.beta0, lane);
        T Tsum = T(0), Wprev = T(0);
#pragma unroll
        for (int i = 0; i <= NM; ++i) {
            if (i > nm) break;
            const double mp = (double)(r0 + i);
            while (ahi <= mp && c < DNU - 1) {
                Tsum += D;
                ++c;
                alo = ahi;
                ahi = A(c + 1);
                Pr0 = Pr1;
                Pr1 = Pv[(long)ucol(min(c + 1, DNU - 1)) * DNV];
                iw = inv_rs * invdk[ucol(c)];
                D = column_delta(g, G, Pr0, p, ucol(c), inv_sj, vp.beta0, lane);
            }
            const double fr = fmin(fmax((mp - alo) * iw, 0.0), 1.0);
            const T W = Tsum + (T)fr * D;
            if (i > 0) acc[i - 1] += W - Wprev;
            Wprev = W;
        }
    }

    if (lane < ZN) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            if (i >= nm) break;
            const int m = r0 + i;
            const long col = yslab ? ((long)j * g.XN + m) : ((long)m * g.XN + j);
            const float mk = (g.mask == nullptr || g.mask[col]) ? 1.0f : 0.0f;
            const long idx = col * ZN + lane;
            const float v = (float)acc[i] * mk;
            vol[idx] = accumulate ? vol[idx] + v : v;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
constexpr int kProjNU = 16;
constexpr int kBackNM = 16;

template <typename T>
hipError_t launch_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s)
{
    T* cz = static_cast<T*>(d_cz);
    const long ncol = (long)g.XN * g.YN;
    hipLaunchKernelGGL(zprefix_kernel<T>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s,
                       d_vol, cz, g.mask, ncol, g.ZN);
    const int ntiles = (g.DNU + kProjNU - 1) / kProjNU;
    hipLaunchKernelGGL((dd_proj_walk<T, kProjNU>), dim3(ntiles, (g.PN + 3) / 4), dim3(256), 0, s,
                       g, cz, d_prj);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_back(const DevGeom& g, const float* d_prj, void* d_P, float* d_vol,
                       const int* d_views_y, int ny, const int* d_views_x, int nx, hipStream_t s)
{
    T* P = static_cast<T*>(d_P);
    const long ncu = (long)g.PN * g.DNU;
    hipLaunchKernelGGL(prjprefix_kernel<T>, dim3((unsigned)((ncu + 255) / 256)), dim3(256), 0, s,
                       g, d_prj, P);
    // pass 1: y-slab views (slab = y, rows along x); overwrites
    {
        const int NS = g.YN, NR = g.XN, ntr = (NR + kBackNM - 1) / kBackNM;
        hipLaunchKernelGGL((dd_back_walk<T, kBackNM>), dim3((unsigned)(ntr * ((NS + 3) / 4))), dim3(256), 0,
                           s, g, P, d_vol, d_views_y, ny, 1, 0, NS, NR, ntr);
    }
    // pass 2: x-slab views (slab = x, rows along y); accumulates
    {
        const int NS = g.XN, NR = g.YN, ntr = (NR + kBackNM - 1) / kBackNM;
        hipLaunchKernelGGL((dd_back_walk<T, kBackNM>), dim3((unsigned)(ntr * ((NS + 3) / 4))), dim3(256), 0,
                           s, g, P, d_vol, d_views_x, nx, 0, 1, NS, NR, ntr);
    }
    return hipGetLastError();
}

template hipError_t launch_proj<double>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_proj<float>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_back<double>(const DevGeom&, const float*, void*, float*, const int*, int,
                                        const int*, int, hipStream_t);
template hipError_t launch_back<float>(const DevGeom&, const float*, void*, float*, const int*, int,
                                       const int*, int, hipStream_t);

}  // namespace ctlab

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

// Intra-wavefront hand-off through LDS (every lane writes its prefix entry,
// then reads two entries written by other lanes): the LDS executes one
// wavefront's instructions in issue order, and the compiler keeps the store /
// load order because the indices may alias, so no fence is needed (a fence
// would also wait for the prefetched global loads still in flight).

// uniform read of a double held by lane l
__device__ __forceinline__ double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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
// inclusive prefix over the 64 lanes of a wavefront (lane l gets sum of lanes 0..l)
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    return v;
}

// One wavefront per voxel column (lane = z; ZN <= 64): coalesced read of the
// column, wavefront scan, coalesced write.
template <typename T>
__global__ void __launch_bounds__(256)
zprefix_kernel(const float* __restrict__ vol, T* __restrict__ cz,
               const unsigned char* __restrict__ mask, long ncol, int ZN)
{
    const int lane = threadIdx.x & 63;
    const long col = blockIdx.x * 4L + (threadIdx.x >> 6);
    if (col >= ncol) return;
    const T mk = (mask == nullptr || mask[col]) ? T(1) : T(0);
    const T v = lane < ZN ? (T)vol[col * ZN + lane] * mk : T(0);
    const T r = wave_inclusive_scan(v, lane);
    if (lane < ZN) cz[col * ZN + lane] = r;
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
    const int u0 = blockIdx.x * NU;
    const int nu = min(NU, g.DNU - u0);
    // this tile's boundary slopes (lanes 0..nu) and magnifications (lanes 0..nu-1)
    // are the same for every slab: load once, read per column with v_readlane
    const double kap_l = g.kap[(long)p * (g.DNU + 1) + u0 + min(lane, nu)];
    const double tau_l = g.tau[(long)p * g.DNU + u0 + min(lane, nu - 1)];
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
        const double afirst = vp.rbase + rs * readlane_d(kap_l, vp.asc ? 0 : nu);
        const double alast = vp.rbase + rs * readlane_d(kap_l, vp.asc ? nu : 0);
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
            double aq = vp.rbase + rs * readlane_d(kap_l, vp.asc ? q : nu - q);
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
                const double t = g.cone ? sj * readlane_d(tau_l, vp.asc ? q - 1 : nu - q) : 1.0;
                G[lane + 1] = H;
                const T Dhi = lds_interp(G, vp.beta0 + t * zhi, ZN);
                const T Dlo = lds_interp(G, vp.beta0 + t * zlo, ZN);
                acc[q - 1] += (Dhi - Dlo) * qj;
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
    const int lane = threadIdx.x & 63;
    const long idx = blockIdx.x * 4L + (threadIdx.x >> 6);  // p*DNU + u, one wavefront each (DNV <= 64)
    if (idx >= (long)g.PN * g.DNU) return;
    const int DNV = g.DNV;
    T v = T(0);
    if (lane < DNV) {
        const double ezv = g.ez[lane];
        const double L = g.dx * sqrt(g.exy2[idx] + ezv * ezv) * fabs(g.tau[idx]);
        v = (T)((double)prj[idx * DNV + lane] * L);
    }
    const T r = wave_inclusive_scan(v, lane);
    if (lane < DNV) P[idx * DNV + lane] = r;
}

// Adjoint z-part for one detector column: sum_v y*L * (fraction of cell v
// covered by the voxel z-interval [lane, lane+1] mapped onto the detector rows),
// read from the column's v-prefix Pu (lane v holds prefix v+1) through LDS.
// esdz = (cell-centre slab coordinate - source) / row pitch, so that the
// voxel boundary b maps to the fractional row (b - beta0) * esdz / sj - zhat0/dzh.
template <typename T>
__device__ __forceinline__ T column_delta(const DevGeom& g, T* G, T Pu, double esdz, double inv_sj,
                                          double beta0, int lane)
{
    const int DNV = g.DNV;
    G[lane + 1] = Pu;
    double nlo, nhi;
    T sgn = T(1);
    if (g.cone) {
        const double inv = inv_sj * esdz;   // 1 / (t * dzh)
        const double off = -beta0 * inv - g.zhat0 / g.dzh;
        nlo = (double)lane * inv + off;
        nhi = nlo + inv;
        if (g.dzh < 0.0) sgn = T(-1);
    } else {
        nlo = (double)lane;
        nhi = (double)(lane + 1);
    }
    return (lds_interp(G, nhi, DNV) - lds_interp(G, nlo, DNV)) * sgn;
}

// Column record in walk order (increasing row coordinate a):
//   x = slope of the lower boundary, y = slope of the upper boundary,
//   z = 1/|slope step|, w = esdz of the column.
struct ColRec { double klo, khi, invdk, esdz; };

// First walk-order column whose upper boundary lies above row r0 = 8*r8, per
// (view, slab, r8); 0xFFFF when every column lies below r0.
__global__ void __launch_bounds__(256)
back_start_kernel(DevGeom g, const ColRec* __restrict__ rec, unsigned short* __restrict__ start, int SJ, int SR)
{
    const long idx = blockIdx.x * 256L + threadIdx.x;
    if (idx >= (long)g.PN * SJ * SR) return;
    const int r8 = (int)(idx % SR);
    const long pj = idx / SR;
    const int j = (int)(pj % SJ);
    const int p = (int)(pj / SJ);
    const ViewParams vp = g.vp[p];
    const int DNU = g.DNU;
    const double r0 = 8.0 * r8;
    if (j >= vp.NS || r0 >= vp.NR) { start[idx] = 0xFFFF; return; }
    const double rs = (vp.s0 + j * g.dx) / g.dx;
    const ColRec* R = rec + (long)p * DNU;
    auto A = [&](int q) { return vp.rbase + rs * (q < DNU ? R[q].klo : R[DNU - 1].khi); };
    if (A(DNU) <= r0) { start[idx] = 0xFFFF; return; }
    int lo = 1, hi = DNU;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A(mid) > r0) hi = mid; else lo = mid + 1;
    }
    start[idx] = (unsigned short)(lo - 1);
}

// Block = 4 wavefronts = 4 consecutive slabs of one tile of NM rows (their
// detector-column ranges nearly coincide); each wavefront loops over the views
// of one slab orientation; lane = z.  Per view: the start column comes from
// the table above (prefetched one view ahead), then columns are consumed in
// walk order with the next column's record and P row in flight.
template <typename T, int NM>
__global__ void __launch_bounds__(256)
dd_back_walk(DevGeom g, const T* __restrict__ P, float* __restrict__ vol,
             const int* __restrict__ views, int nviews, int yslab, int accumulate,
             int NS, int NR, int ntr, const ColRec* __restrict__ rec,
             const unsigned short* __restrict__ start, int SJ, int SR)
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
    const long srow = (long)j * SR + (r0 >> 3);

    T acc[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] = T(0);

    int pn = nviews > 0 ? views[0] : 0;
    int stn = nviews > 0 ? start[(long)pn * SJ * SR + srow] : 0xFFFF;
    for (int iv = 0; iv < nviews; ++iv) {
        const int p = __builtin_amdgcn_readfirstlane(pn);
        const int st = __builtin_amdgcn_readfirstlane(stn);
        if (iv + 1 < nviews) {
            pn = views[iv + 1];
            stn = start[(long)pn * SJ * SR + srow];
        }
        if (st == 0xFFFF) continue;
        const ViewParams vp = g.vp[p];
        const double sj = vp.s0 + j * g.dx;
        const double rs = sj * inv_dx;
        const double inv_sj = 1.0 / sj;
        const double inv_rs = fabs(g.dx * inv_sj);
        const int asc = vp.asc;
        const ColRec* __restrict__ R = rec + (long)p * DNU;
        const T* __restrict__ Pv = P + (long)p * DNU * DNV + vl;
        int c = st;
        ColRec r = R[c];
        double alo = vp.rbase + rs * r.klo;
        if (alo >= (double)(r0 + nm)) continue;
        double ahi = vp.rbase + rs * r.khi;
        double iw = inv_rs * r.invdk;
        const int c1 = min(c + 1, DNU - 1);
        T Pr0 = Pv[(long)(asc ? c : DNU - 1 - c) * DNV];
        ColRec rn = R[c1];
        T Pr1 = Pv[(long)(asc ? c1 : DNU - 1 - c1) * DNV];
        T D = column_delta(g, G, Pr0, r.esdz, inv_sj, vp.beta0, lane);
        T Tsum = T(0), Wprev = T(0);
#pragma unroll
        for (int i = 0; i <= NM; ++i) {
            if (i > nm) break;
            const double mp = (double)(r0 + i);
            while (ahi <= mp && c < DNU - 1) {
                Tsum += D;
                ++c;
                alo = ahi;
                ahi = vp.rbase + rs * rn.khi;
                iw = inv_rs * rn.invdk;
                D = column_delta(g, G, Pr1, rn.esdz, inv_sj, vp.beta0, lane);
                const int cn = min(c + 1, DNU - 1);
                rn = R[cn];
                Pr1 = Pv[(long)(asc ? cn : DNU - 1 - cn) * DNV];
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
template <typename T, int NU>
static void launch_proj_walk(const DevGeom& g, const T* cz, float* d_prj, hipStream_t s)
{
    const int ntiles = (g.DNU + NU - 1) / NU;
    hipLaunchKernelGGL((dd_proj_walk<T, NU>), dim3(ntiles, (g.PN + 3) / 4), dim3(256), 0, s, g, cz, d_prj);
}

template <typename T, int NM>
static void launch_back_walk(const DevGeom& g, const T* P, float* d_vol, const int* views, int nv, int yslab,
                             int acc, const ColRec* rec, const unsigned short* start, int SJ, int SR, hipStream_t s)
{
    const int NS = yslab ? g.YN : g.XN, NR = yslab ? g.XN : g.YN, ntr = (NR + NM - 1) / NM;
    hipLaunchKernelGGL((dd_back_walk<T, NM>), dim3((unsigned)(ntr * ((NS + 3) / 4))), dim3(256), 0, s, g, P, d_vol,
                       views, nv, yslab, acc, NS, NR, ntr, rec, start, SJ, SR);
}

template <typename T>
hipError_t launch_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s)
{
    T* cz = static_cast<T*>(d_cz);
    const long ncol = (long)g.XN * g.YN;
    hipLaunchKernelGGL(zprefix_kernel<T>, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, s,
                       d_vol, cz, g.mask, ncol, g.ZN);
    if (g.tile_nu == 32) launch_proj_walk<T, 32>(g, cz, d_prj, s);
    else if (g.tile_nu == 8) launch_proj_walk<T, 8>(g, cz, d_prj, s);
    else launch_proj_walk<T, 16>(g, cz, d_prj, s);
    return hipGetLastError();
}

hipError_t launch_back_start(const DevGeom& g, const void* d_rec, unsigned short* d_start, int SJ, int SR,
                             hipStream_t s)
{
    const long n = (long)g.PN * SJ * SR;
    hipLaunchKernelGGL(back_start_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g,
                       (const ColRec*)d_rec, d_start, SJ, SR);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_back(const DevGeom& g, const float* d_prj, void* d_P, float* d_vol,
                       const int* d_views_y, int ny, const int* d_views_x, int nx, const void* d_rec,
                       const unsigned short* d_start, int SJ, int SR, hipStream_t s)
{
    const ColRec* rec = (const ColRec*)d_rec;
    T* P = static_cast<T*>(d_P);
    const long ncu = (long)g.PN * g.DNU;
    hipLaunchKernelGGL(prjprefix_kernel<T>, dim3((unsigned)((ncu + 3) / 4)), dim3(256), 0, s,
                       g, d_prj, P);
    // pass 1: y-slab views (slab = y, rows along x) overwrite; pass 2: x-slab views accumulate
    for (int pass = 0; pass < 2; ++pass) {
        const int ys = pass == 0;
        const int* vw = ys ? d_views_y : d_views_x;
        const int nv = ys ? ny : nx;
        if (g.tile_nm == 32) launch_back_walk<T, 32>(g, P, d_vol, vw, nv, ys, pass, rec, d_start, SJ, SR, s);
        else if (g.tile_nm == 8) launch_back_walk<T, 8>(g, P, d_vol, vw, nv, ys, pass, rec, d_start, SJ, SR, s);
        else launch_back_walk<T, 16>(g, P, d_vol, vw, nv, ys, pass, rec, d_start, SJ, SR, s);
    }
    return hipGetLastError();
}

template hipError_t launch_proj<double>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_proj<float>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_back<double>(const DevGeom&, const float*, void*, float*, const int*, int,
                                        const int*, int, const void*, const unsigned short*, int, int, hipStream_t);
template hipError_t launch_back<float>(const DevGeom&, const float*, void*, float*, const int*, int,
                                       const int*, int, const void*, const unsigned short*, int, int, hipStream_t);

}  // namespace ctlab

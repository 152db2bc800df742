// Fused elementwise / reduction kernels of the reconstruction loops (gfx950).
//
//   SART_Weighting<T>::Proj / Back ...... src/SARTWeighting.h:22-136 (eps 1e-7)
//   ossart_update_functor ............... src/utilities.hpp:1202-1223 (IS_ZERO 1e-9)
//   FISTA_functor / FISTA ............... src/FastIterativeShrinkageAlgorithm.h:17-69
//   _softThreshold_functor .............. src/utilities.hpp:1097-1117
//   CG_update{1,2,3}_functor ............ src/utilities.hpp:1121-1156
//   _betaupdate_functor + calculateBetaForCG  src/utilities.hpp:1228-1268, 1611-1621
// Each reference thrust::transform becomes one grid-stride kernel (HBM-bound,
// float4 where the sizes allow); every reduction accumulates in FP64 and ends
// in a fixed-order second pass, so results do not depend on the launch shape.
// The TV proximal operator (FGP, Beck & Teboulle 2009) has no reference body.
#include <hip/hip_runtime.h>

#include "status.hpp"

namespace ctlab {

constexpr int kEwBlock = 256;
constexpr int kRedBlocks = 1024;

static inline unsigned ew_grid(long n)
{
    long b = (n + kEwBlock - 1) / kEwBlock;
    if (b > 65536L * 8) b = 65536L * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

#define GRID_STRIDE(i, n) \
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// prj = rowSum > 1e-7 ? (real - prj) / rowSum : 0
__global__ void k_sart_prj_weight(float* __restrict__ prj, const float* __restrict__ real,
                                  const float* __restrict__ rowSum, long n)
{
    GRID_STRIDE(i, n)
    {
        const float w = rowSum[i];
        prj[i] = w > 1.0e-7f ? (real[i] - prj[i]) / w : 0.0f;
    }
}

// recon += colSum > 1e-7 ? lambda * vol / colSum : 0, then optional clamp at lo
__global__ void k_sart_bak_weight(const float* __restrict__ vol, float* __restrict__ recon,
                                  const float* __restrict__ colSum, float lambda, int clamp_lo, float lo, long n)
{
    GRID_STRIDE(i, n)
    {
        const float w = colSum[i];
        float r = recon[i] + (w > 1.0e-7f ? lambda * (vol[i] / w) : 0.0f);
        if (clamp_lo && r < lo) r = lo;
        recon[i] = r;
    }
}

// img = |weg| >= 1e-9 ? (img + lambda*cor/weg)*msk : 0; IS_ZERO compares the
// float against the double constant 1e-9 (src/utilities.hpp:316-320)
__global__ void k_ossart_update(float* __restrict__ img, const float* __restrict__ cor,
                                const float* __restrict__ weg, const float* __restrict__ msk, float lambda, long n)
{
    GRID_STRIDE(i, n)
    {
        const float w = weg[i];
        const bool zero = ((double)w < 1.0e-9) && ((double)w > -1.0e-9);
        img[i] = zero ? 0.0f : (img[i] + lambda * cor[i] / w) * msk[i];
    }
}

// cur = clamp(cur + (t1-1)/t2 * (cur - las), minV, maxV), with the reference
// functor's arithmetic: (t1 - 1.0)/t2 and the sum in double, (cur - las) in
// float, result rounded to float before the clamp
__global__ void k_fista(float* __restrict__ cur, const float* __restrict__ las, float t1, float t2,
                        float minV, float maxV, long n)
{
    const double c = ((double)t1 - 1.0) / (double)t2;
    GRID_STRIDE(i, n)
    {
        const float x = cur[i];
        const float r = (float)((double)x + c * (double)(x - las[i]));
        cur[i] = r < minV ? minV : (r > maxV ? maxV : r);
    }
}

__global__ void k_soft_threshold(float* __restrict__ x, float g, long n)
{
    GRID_STRIDE(i, n)
    {
        const float r = x[i];
        x[i] = r < -g ? r + g : (r > g ? r - g : 0.0f);
    }
}

// y = a*x + b*y  (covers CG_update1: X + alpha*D, update2: R - alpha*tem2, update3: R + beta*D)
__global__ void k_axpby(float a, const float* __restrict__ x, float b, float* __restrict__ y, long n)
{
    GRID_STRIDE(i, n) y[i] = a * x[i] + b * y[i];
}

// out = x + a*(y - z)   (gradient / residual combinations)
__global__ void k_xpaymz(float* __restrict__ out, const float* __restrict__ x, float a,
                         const float* __restrict__ y, const float* __restrict__ z, long n)
{
    GRID_STRIDE(i, n) out[i] = x[i] + a * (y[i] - (z ? z[i] : 0.0f));
}

__global__ void k_fill(float* __restrict__ x, float v, long n)
{
    GRID_STRIDE(i, n) x[i] = v;
}

__device__ __forceinline__ double warp_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block partials of two FP64 sums: case semantics of _betaupdate_functor with
// (gk, gk1, d) = (-r0, -r1, d) when from_residual is set (the loop keeps residuals)
__global__ void __launch_bounds__(256)
k_beta_partial(const float* __restrict__ r0, const float* __restrict__ r1, const float* __restrict__ d,
               int cases, int from_residual, long n, double* __restrict__ part)
{
    double s0 = 0.0, s1 = 0.0;
    const double sg = from_residual ? -1.0 : 1.0;
    GRID_STRIDE(i, n)
    {
        const double gk = sg * r0[i], gk1 = sg * r1[i], e = d[i];
        switch (cases) {
        case 1: s0 += gk1 * (gk1 - gk); s1 += gk * gk; break;
        case 2: s0 += gk1 * (gk1 - gk); s1 += e * (gk1 - gk); break;
        case 3: s0 += gk1 * gk1; s1 += -e * gk; break;
        case 4: s0 += gk1 * gk1; s1 += e * (gk1 - gk); break;
        default: s0 += gk1 * gk1; s1 += gk * gk; break;
        }
    }
    __shared__ double sh[2][4];
    s0 = warp_sum(s0);
    s1 = warp_sum(s1);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[0][w] = s0; sh[1][w] = s1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

// block partials of <x,y> (y == nullptr: <x,x>) and, optionally, <x,z>
__global__ void __launch_bounds__(256)
k_dot_partial(const float* __restrict__ x, const float* __restrict__ y, long n, double* __restrict__ part)
{
    double s = 0.0;
    GRID_STRIDE(i, n)
    {
        const double a = x[i];
        s += a * (y ? (double)y[i] : a);
    }
    __shared__ double sh[4];
    s = warp_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// fixed-order final sum of `nb` partials with stride `stride` (one block)
__global__ void __launch_bounds__(256)
k_sum_partials(const double* __restrict__ part, int nb, int stride, int nsums, double* __restrict__ out)
{
    __shared__ double sh[4];
    for (int k = 0; k < nsums; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)b * stride + k];
        s = warp_sum(s);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) out[k] = sh[0] + sh[1] + sh[2] + sh[3];
        __syncthreads();
    }
}

// ---- FGP TV prox on a (Z,X,Y) volume: p holds 3 components of n each ------
// Launch: grid (ceil(XN*ZN / 256), YN), one thread per voxel of row y; the
// in-row index j = x*ZN + z splits by a float reciprocal of ZN and one
// correction step (exact for j < 2^22 * ZN: the estimate is off by at most 1),
// instead of 64-bit divisions of the flat index.
struct RowIdx { int x, z; long i; bool ok; };
__device__ __forceinline__ RowIdx row_index(int XN, int ZN, float rcpZN)
{
    RowIdx r;
    const int j = blockIdx.x * 256 + threadIdx.x;
    r.ok = j < XN * ZN;
    int q = (int)((float)j * rcpZN);
    int z = j - q * ZN;
    if (z < 0) { --q; z += ZN; } else if (z >= ZN) { ++q; z -= ZN; }
    r.x = q;
    r.z = z;
    r.i = (long)blockIdx.y * XN * ZN + j;
    return r;
}

// tmp = b - lam * L(r); L(r) = sum_d r_d[i] - r_d[i - e_d] (0 outside / at the far face)
__global__ void __launch_bounds__(256) k_tv_primal(const float* __restrict__ b, const float* __restrict__ r, float lam,
                                                   int XN, int YN, int ZN, float rcpZN, float* __restrict__ out)
{
    const long n = (long)XN * YN * ZN;
    const RowIdx q = row_index(XN, ZN, rcpZN);
    if (!q.ok) return;
    const long i = q.i;
    const float* rx = r;
    const float* ry = r + n;
    const float* rz = r + 2 * n;
    float v = rx[i] + ry[i] + rz[i];
    if (q.x > 0) v -= rx[i - ZN];
    if (blockIdx.y > 0) v -= ry[i - (long)XN * ZN];
    if (q.z > 0) v -= rz[i - 1];
    out[i] = b[i] - lam * v;
}

// One FGP iteration, fused.  The extrapolated dual r = pc + mom*(pc - pp) (pc,
// pp: the last two dual iterates, 3n floats each) is formed where it is read;
// tmp = b - lam * L(r) (k_tv_primal's arithmetic) is needed at the voxel and
// its three forward neighbours; then
//   pn = proj_{|.|<=1}(r + step * L^T(tmp))
// goes to a third buffer, so no thread writes what another one reads.  Per
// iteration this moves 9 floats per voxel through HBM (pc, pp, b, pn) where a
// primal and a dual pass over stored r and tmp moved 17.
// A block covers 256 consecutive points of one y-plane (z fastest, then x):
// it forms tmp once at each of them and at the ZN points after them (the x + 1
// neighbours of its last row) into LDS, and reads the x + 1 and z + 1
// neighbours' tmp from there; only the y + 1 neighbour's is formed per thread.
// Launched with (256 + ZN) floats of dynamic LDS.
// R: kDualAll reads pc and pp; kDualNoMom (momentum exactly 0, the second
// iteration) reads pc only, r = pc; kDualZero (the first iteration, from
// p = 0) reads neither, r = 0.  Each equals kDualAll on the zero-filled
// buffers bit for bit, without the fill.
enum { kDualAll = 0, kDualNoMom = 1, kDualZero = 2 };
template <int R>
__global__ void __launch_bounds__(256) k_tv_fgp(const float* __restrict__ b, const float* __restrict__ pc,
                                                const float* __restrict__ pp, float mom, float lam, float step,
                                                int XN, int YN, int ZN, float rcpZN, float* __restrict__ pn)
{
    extern __shared__ float T[];
    const long n = (long)XN * YN * ZN;
    const int y = blockIdx.y;
    const long ex = ZN, ey = (long)XN * ZN;
    const int nplane = XN * ZN;
    const int j0 = blockIdx.x * 256;
    auto r = [&](int c, long j) {
        if constexpr (R == kDualZero) return 0.0f;
        const float a = pc[c * n + j];
        if constexpr (R == kDualNoMom) return a;
        return a + mom * (a - pp[c * n + j]);
    };
    // Every load is unconditional: a neighbour that does not exist is read at
    // a valid index and its term masked by a select.  Subtracting an exact 0
    // leaves each sum bit for bit as k_tv_primal forms it.
    auto tmp_at = [&](long j, bool mx, long jx, bool my, long jy, bool mz, long jz) {
        float v = r(0, j) + r(1, j) + r(2, j);
        const float sx = r(0, jx), sy = r(1, jy), sz = r(2, jz);
        v -= mx ? sx : 0.0f;
        v -= my ? sy : 0.0f;
        v -= mz ? sz : 0.0f;
        return b[j] - lam * v;
    };
    const bool ly = y > 0, hy = y < YN - 1;
    for (int k = threadIdx.x; k < 256 + ZN; k += 256) {
        const int j = j0 + k;
        if (j < nplane) {
            int xq = (int)((float)j * rcpZN);
            int zq = j - xq * ZN;
            if (zq < 0) { --xq; zq += ZN; } else if (zq >= ZN) { ++xq; zq -= ZN; }
            const long i = (long)y * ey + j;
            const bool mx = xq > 0, mz = zq > 0;
            T[k] = tmp_at(i, mx, mx ? i - ex : i, ly, ly ? i - ey : i, mz, mz ? i - 1 : i);
        }
    }
    __syncthreads();
    const RowIdx q = row_index(XN, ZN, rcpZN);
    if (!q.ok) return;
    const long i = q.i;
    const bool lx = q.x > 0, lz = q.z > 0;                               // backward neighbours
    const bool hx = q.x < XN - 1, hz = q.z < ZN - 1;                     // forward neighbours
    auto nb = [&](bool ok, long j) { return ok ? j : i; };
    const float t = T[threadIdx.x];
    const float tX = T[threadIdx.x + ZN];   // read only where hx (inside the plane)
    const float tZ = T[threadIdx.x + 1];    // read only where hz (same x row)
    const long jY = nb(hy, i + ey);
    const float tY = tmp_at(jY, lx, nb(hy && lx, i + ey - ex), true, i, lz, nb(hy && lz, i + ey - 1));
    const float qx = hx ? t - tX : 0.0f;
    const float qy = hy ? t - tY : 0.0f;
    const float qz = hz ? t - tZ : 0.0f;
    const float a0 = r(0, i) + step * qx, a1 = r(1, i) + step * qy, a2 = r(2, i) + step * qz;
    const float nrm = sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
    const float sc = nrm > 1.0f ? 1.0f / nrm : 1.0f;
    pn[i] = a0 * sc;
    pn[n + i] = a1 * sc;
    pn[2 * n + i] = a2 * sc;
}

// ---------------------------------------------------------------------------
// launchers (stream-ordered; no allocation, no memset: capturable in graphs)
// ---------------------------------------------------------------------------
hipError_t rk_sart_prj_weight(float* prj, const float* real, const float* rowSum, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_sart_prj_weight, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, prj, real, rowSum, n);
    return hipGetLastError();
}

hipError_t rk_sart_bak_weight(const float* vol, float* recon, const float* colSum, float lambda, int clamp_lo,
                              float lo, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_sart_bak_weight, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, vol, recon, colSum, lambda,
                       clamp_lo, lo, n);
    return hipGetLastError();
}

hipError_t rk_ossart_update(float* img, const float* cor, const float* weg, const float* msk, float lambda,
                            long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_ossart_update, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, img, cor, weg, msk, lambda, n);
    return hipGetLastError();
}

hipError_t rk_fista(float* cur, const float* las, float t1, float t2, float minV, float maxV, long n,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_fista, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, cur, las, t1, t2, minV, maxV, n);
    return hipGetLastError();
}

hipError_t rk_soft_threshold(float* x, float g, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_soft_threshold, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, x, g, n);
    return hipGetLastError();
}

hipError_t rk_axpby(float a, const float* x, float b, float* y, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, a, x, b, y, n);
    return hipGetLastError();
}

hipError_t rk_xpaymz(float* out, const float* x, float a, const float* y, const float* z, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_xpaymz, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, out, x, a, y, z, n);
    return hipGetLastError();
}

hipError_t rk_fill(float* x, float v, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(ew_grid(n)), dim3(kEwBlock), 0, s, x, v, n);
    return hipGetLastError();
}

// part: >= 2*kRedBlocks doubles of scratch; out: 2 doubles (num, den)
hipError_t rk_beta(const float* r0, const float* r1, const float* d, int cases, int from_residual, long n,
                   double* part, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_beta_partial, dim3(kRedBlocks), dim3(256), 0, s, r0, r1, d, cases, from_residual, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, kRedBlocks, 2, 2, out);
    return hipGetLastError();
}

// out[0] = <x, y> (y null: <x, x>)
hipError_t rk_dot(const float* x, const float* y, long n, double* part, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, s, x, y, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, kRedBlocks, 1, 1, out);
    return hipGetLastError();
}

int red_scratch_doubles() { return 2 * kRedBlocks; }

// FGP prox: x = argmin 0.5||x-b||^2 + lam TV(x); work: 9n floats (three dual
// iterates of 3n, rotated).  Step 1/(4*ndim*lam); momentum from the FISTA
// t-sequence (host computed).
hipError_t rk_tv_prox(const float* b, float* x, float lam, int iters, int XN, int YN, int ZN, float* work,
                      hipStream_t s)
{
    const long n = (long)XN * YN * ZN;
    if (lam <= 0.0f) {
        return hipMemcpyAsync(x, b, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    }
    // row index and grid limits; the FGP kernel's (256 + ZN) floats of LDS within 64 KiB
    if ((long)XN * ZN > (1L << 22) || YN > 65535 || ZN > 16128) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(((long)XN * ZN + 255) / 256), (unsigned)YN);
    const float rcp = 1.0f / (float)ZN;
    const int nd = (XN > 1) + (YN > 1) + (ZN > 1);
    const float step = 1.0f / (4.0f * (float)(nd > 0 ? nd : 1) * lam);
    float* P[3] = {work, work + 3 * n, work + 6 * n};
    // p_0 = 0 and the iterate before it = 0: the first r is 0, which the first
    // iteration uses without reading the buffers; with no iteration the
    // primal reads p_0 itself
    if (iters <= 0) hipLaunchKernelGGL(k_fill, dim3(ew_grid(3 * n)), dim3(kEwBlock), 0, s, work, 0.0f, 3 * n);
    int cur = 0, prev = 1;
    float mom = 0.0f;
    double t = 1.0;
    const size_t lds = (256 + ZN) * sizeof(float);
    for (int it = 0; it < iters; ++it) {
        const int nxt = 3 - cur - prev;
        if (it == 0)
            hipLaunchKernelGGL(k_tv_fgp<kDualZero>, grid, dim3(256), lds, s, b, P[cur], P[prev], mom, lam, step, XN,
                               YN, ZN, rcp, P[nxt]);
        else if (mom == 0.0f)
            hipLaunchKernelGGL(k_tv_fgp<kDualNoMom>, grid, dim3(256), lds, s, b, P[cur], P[prev], mom, lam, step, XN,
                               YN, ZN, rcp, P[nxt]);
        else
            hipLaunchKernelGGL(k_tv_fgp<kDualAll>, grid, dim3(256), lds, s, b, P[cur], P[prev], mom, lam, step, XN,
                               YN, ZN, rcp, P[nxt]);
        const double tn = 0.5 * (1.0 + __builtin_sqrt(1.0 + 4.0 * t * t));
        mom = (float)((t - 1.0) / tn);                  // extrapolates from this iteration's output
        t = tn;
        prev = cur;
        cur = nxt;
    }
    hipLaunchKernelGGL(k_tv_primal, grid, dim3(256), 0, s, b, P[cur], lam, XN, YN, ZN, rcp, x);
    return hipGetLastError();
}

}  // namespace ctlab

// Per-wavefront bodies of the branchless DD kernels, shared by the per-plan
// kernels (dd_kernels.hip) and the batched persistent kernels (batch.hip).
#pragma once
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

__device__ __forceinline__ float readlane_t(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double readlane_t(double v, int l) { return readlane_d(v, l); }

// lane i receives lane i+1's value (DPP wave_shl:1, a VALU move); lane 63 receives 0
__device__ __forceinline__ double shift_down1(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float shift_down1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// lane i receives lane i+1's value; lane 63, which has no source lane, keeps
// `fill` (DPP wave_shl:1 without bound_ctrl leaves that lane unwritten)
__device__ __forceinline__ double shift_down1_fill(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = __builtin_amdgcn_update_dpp((int)(f & 0xffffffffLL), (int)(b & 0xffffffffLL), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(f >> 32), (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float shift_down1_fill(float v, float fill)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// clamp i to [0, hi] in one VALU instruction (hi uniform)
__device__ __forceinline__ int clamp_idx(int i, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(i), "s"(hi));
    return r;
}

// piecewise-linear read of a prefix table G[0..n] (G[0] = 0) at b, clamped
// to [0, n]: the cell index is clamped as an integer and the fraction to
// [0, 1], which gives G[0] below the table and G[n] above it
template <typename T>
__device__ __forceinline__ T lds_interp(const T* G, double b, int n)
{
    const int i = clamp_idx((int)b, n - 1);
    const T fr = (T)fmin(fmax(b - (double)i, 0.0), 1.0);
    const T g0 = G[i];
    const T g1 = G[i + 1];
    return g0 + fr * (g1 - g0);
}

// ---------------------------------------------------------------------------
// z-prefix of every masked voxel column: cz[col*ZN + k] = sum_{k'<=k} f*mask
// ---------------------------------------------------------------------------
// DPP move of v with control `ctrl` on the rows in row_mask; lanes without a
// source, and rows outside the mask, receive 0
template <int ctrl, int row_mask>
__device__ __forceinline__ double dpp_or0(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), ctrl, row_mask, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, row_mask, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int ctrl, int row_mask>
__device__ __forceinline__ float dpp_or0(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, row_mask, 0xf, true));
}

// inclusive prefix over the 64 lanes of a wavefront (lane l gets sum of lanes
// 0..l), in DPP moves: shifts by 1, 2, 4, 8 lanes within each row of 16, then
// the last lane of row 0 broadcast into row 1 (and of row 2 into row 3), then
// lane 31 into rows 2 and 3
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    v += dpp_or0<0x111, 0xf>(v);   // row_shr:1
    v += dpp_or0<0x112, 0xf>(v);   // row_shr:2
    v += dpp_or0<0x114, 0xf>(v);   // row_shr:4
    v += dpp_or0<0x118, 0xf>(v);   // row_shr:8
    v += dpp_or0<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_or0<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

// inclusive z-prefix of one masked voxel column (lane = z; ZN <= 64)
template <typename T>
__device__ __forceinline__ void zprefix_wave(const CTLAB_GMEM float* __restrict__ vol, CTLAB_GMEM T* __restrict__ cz,
                                             const CTLAB_GMEM unsigned char* __restrict__ mask, long col, int ZN, int lane)
{
    const T mk = (mask == nullptr || mask[col]) ? T(1) : T(0);
    const T v = lane < ZN ? (T)vol[col * ZN + lane] * mk : T(0);
    const T r = wave_inclusive_scan(v);
    if (lane < ZN) cz[col * ZN + lane] = r;
}

// buffer-resource word 3 for raw (untyped, unswizzled) buffer access on gfx9
constexpr int kBufWord3 = 0x00020000;

// raw buffer load of one element at byte offset off; an offset at or past the
// resource's size reads 0
template <typename T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return __longlong_as_double(((long long)b.y << 32) | b.x);
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// LDS per projector wavefront: one z-prefix table (entries 0..64) per column
// step of the tile, kGStride entries apart
constexpr int kGStride = 68;

// Projector walk of one (view p, tile of NU detector columns) over all slabs.
// Lane = z for the SAT rows, lane = detector row boundary for the outputs;
// G: NU * kGStride entries of LDS, FT: 2 * (NU + 1) doubles of LDS (16-byte
// aligned).
//
// Row v needs the footprint z-prefix at its two boundaries v and v+1: each lane
// interpolates boundary `lane`.  In FP64 the interpolated prefixes are summed
// over slabs per boundary and differenced once at the end (the sum is linear,
// and FP64 keeps the cancellation far below the FP32 output); the FP32 variant
// differences every slab's term before summing.  With 64 rows the top
// boundary (64) has no lane: its table stays in LDS (one region per column
// step) and after the slab's column loop one pass with lane = step
// interpolates all of them; row 63 collects that term lane-parallel in
// acc_top and adds it in the epilogue.
//
// Per slab the boundary fractions and column magnifications go to the step
// table FT as pairs in walk order (step k: fraction of the boundary it ends
// at, magnification of the column it crosses), so that every step reads its
// two uniform operands with one broadcast LDS read.
//
// Rows are read two ahead through a pointer that steps one row at a time; a
// read-ahead past the walk's last row in the slab is issued with a buffer
// resource of size 0 (a scalar select of its size, no branch: a branch around
// each read measured 18 % slower, it breaks the load pipelining), so it moves
// no data: re-reading the last row instead cost 2 % of the projector time in
// L1 traffic.  Only the slabs the tile can reach are walked, an interval of
// slabs found once per tile.
template <typename T, int NU, bool CONE>
__device__ __forceinline__ void proj_walk_wave(const DevGeom& g, const CTLAB_GMEM T* __restrict__ cz,
                                               CTLAB_GMEM float* __restrict__ prj,
                                               int p, int tile, int lane, T* G, double* FT)
{
    constexpr bool kSumThenDiff = sizeof(T) == 8;
    if (lane < NU) G[lane * kGStride] = T(0);
    const ViewParams vp = g.vp[p];
    const int u0 = tile * NU;
    const int nu = min(NU, g.DNU - u0);
    const int ZN = g.ZN, NR = vp.NR, DNV = g.DNV;
    // this tile's boundary slopes (lanes 0..nu) and magnifications, by column
    // (lanes 0..nu-1) and by column step (tau_w), are the same for every slab
    const double kap_l = g.kap[(long)p * (g.DNU + 1) + u0 + min(lane, nu)];
    const double tau_l = g.tau[(long)p * g.DNU + u0 + min(lane, nu - 1)];
    const int uw = min(lane, nu - 1);
    const double tau_w = g.tau[(long)p * g.DNU + u0 + (vp.asc ? uw : nu - 1 - uw)];
    // step-table slots: column `lane` is crossed at step kt, boundary `lane`
    // ends step kf; lanes without a step write the spare slot NU
    const int kt = lane < nu ? (vp.asc ? lane : nu - 1 - lane) : NU;
    const int kf0 = vp.asc ? lane - 1 : nu - 1 - lane;
    const int kf = (lane <= nu && kf0 >= 0) ? kf0 : NU;
    const bool defer_top = DNV == 64;
    // byte offset of lane z within a row (rows are read as uniform row pointer + 32-bit lane offset)
    const unsigned zoff = (unsigned)min(lane, ZN - 1) * (unsigned)sizeof(T);
    const double zb = g.zhat[min(lane, DNV)];        // boundary `lane`
    const double ztop = g.zhat[DNV];
    const double inv_dx = 1.0 / g.dx;
    const long slab_off = (long)vp.slab_stride * ZN;
    const long row_off = (long)vp.row_stride * ZN;
    T acc[NU];
#pragma unroll
    for (int q = 0; q < NU; ++q) acc[q] = T(0);
    T acc_top = T(0);   // lane s: row 63's top-boundary term of column step s

    const CTLAB_GMEM double* __restrict__ invs = g.invs + (long)p * g.sjs;
    // first and last tile boundary in walk order (increasing row coordinate)
    const int b0 = vp.asc ? 0 : nu, bl = vp.asc ? nu : 0;
    // Slabs the tile can reach: boundary l lies at row coordinate
    // a_l(j) = rbase + (s0/dx + j) * kap_l, linear in j, and slab j is needed
    // while a_bl(j) > 0 and a_b0(j) < NR.  Each condition bounds j on one side;
    // the interval is widened by one slab each way against rounding (a slab
    // the tile misses has every boundary clamped to one end of the slab and
    // no active column, so it adds nothing).
    int j_lo = 0, j_hi = vp.NS - 1;
    {
        const double r0 = vp.s0 * inv_dx;
        auto bound = [&](double k, double lim, bool above) {
            // j with rbase + (r0 + j) * k > lim (above) or < lim (below)
            if (k == 0.0) {
                if (above ? !(vp.rbase > lim) : !(vp.rbase < lim)) j_hi = -1;
                return;
            }
            // the condition holds for j > jt (first slab floor(jt) + 1) or for
            // j < jt (last slab ceil(jt) - 1); one slab wider each way
            const double jt = (lim - vp.rbase) / k - r0;
            const double lim_j = (double)vp.NS;
            if ((k > 0.0) == above) j_lo = max(j_lo, (int)fmin(fmax(floor(jt), -1.0), lim_j));
            else j_hi = min(j_hi, (int)fmin(fmax(ceil(jt), -1.0), lim_j));
        };
        bound(readlane_d(kap_l, bl), 0.0, true);
        bound(readlane_d(kap_l, b0), (double)NR, false);
    }
    if constexpr (!CONE) FT[2 * kt + 1] = 1.0;   // multi-slice fan beam: no z magnification
    // slab j's centre relative to the source, and every boundary's row
    // coordinate at slab j (lane = boundary), both advanced one slab at a time
    double sj = vp.s0 + j_lo * g.dx;
    double a_raw = vp.rbase + (vp.s0 * inv_dx + j_lo) * kap_l;
    for (int j = j_lo; j <= j_hi; ++j, sj += g.dx, a_raw += kap_l) {
        // row coordinate clamped to the slab; its row index and fraction within the row
        const double a_l = fmin(fmax(a_raw, 0.0), (double)NR);
        const int m_l = min((int)a_l, NR - 1);
        const double f_l = a_l - (double)m_l;
        FT[2 * kf] = f_l;
        if constexpr (CONE) FT[2 * kt + 1] = sj * tau_l;   // z magnification of column `lane`
        // columns (lane = tile column u, boundaries u and u+1) with a positive
        // row extent: the unclamped coordinates differ, so the clamped ones
        // differ where the column overlaps the slab
        const double a_n = shift_down1(a_l);
        const unsigned act = (unsigned)__ballot(lane < nu && a_n != a_l);
        // the same bits in walk order: bit k = the column crossed at step k
        const unsigned actk = vp.asc ? act : __builtin_bitreverse32(act) >> (32 - nu);
        const double inv_sj = invs[j];
        const T qj = (T)(CONE ? inv_sj * inv_sj : fabs(inv_sj));
        int m = __builtin_amdgcn_readlane(m_l, b0);
        // rows m and m+1 in flight; rp: next row to read (uniform)
        const CTLAB_GMEM T* __restrict__ rp = cz + j * slab_off + (long)m * row_off;
        // a row is read through a buffer resource built from the uniform row
        // pointer (scalar instructions only) with the lane's offset in z; a
        // read-ahead past the walk's last row in the slab gets a resource of
        // size 0, which reads nothing (and returns 0, never used)
        auto row = [zoff, ZN](const CTLAB_GMEM T* r, bool live) {
            const __amdgpu_buffer_rsrc_t rr =
                __builtin_amdgcn_make_buffer_rsrc((void*)r, (short)0, live ? ZN * (int)sizeof(T) : 0, kBufWord3);
            return buf_load<T>(rr, zoff);
        };
        const int m_end = __builtin_amdgcn_readlane(m_l, bl);   // the walk's last row in the slab
        int mr = m + 2;   // row index of rp
        T c0 = row(rp, true);
        T c1 = row(rp + row_off, m + 1 <= m_end);
        rp += 2 * row_off;
        T S = T(0);
        T Eprev = (T)readlane_d(f_l, b0) * c0;
        // every tile runs NU steps: in a partial tile (nu < NU) the steps past
        // nu stay on the last boundary and have no active column
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            const int bq = vp.asc ? min(k + 1, nu) : max(nu - 1 - k, 0);   // boundary reached at this step
            const int mq = __builtin_amdgcn_readlane(m_l, bq);
            while (m < mq) {
                S += c0;
                ++m;
                c0 = c1;
                c1 = row(rp, mr <= m_end);
                rp += row_off;
                ++mr;
            }
            // both step operands in one LDS read: the empty asm makes the
            // magnification, used only in the active branch, live here
            const double2 ft = *reinterpret_cast<const double2*>(FT + 2 * k);
            asm volatile("" ::"v"(ft.y));
            const T E = S + (T)ft.x * c0;
            const T H = E - Eprev;
            Eprev = E;
            if ((actk >> k) & 1u) {
                T* Gq = G + k * kGStride;
                Gq[lane + 1] = H;
                const T Db = lds_interp(Gq, fma(ft.y, zb, vp.beta0), ZN);
                if constexpr (kSumThenDiff) {
                    acc[k] += Db * qj;
                } else {
                    const T Dn = shift_down1(Db);   // lane 63: 0, its top boundary is deferred
                    acc[k] += (Dn - Db) * qj;
                }
            }
        }
        if (defer_top) {
            // lane s < nu: top boundary of column step s (its table is region s);
            // the lane test keeps the shift below 32 bits
            if (lane < nu && ((actk >> lane) & 1u)) {
                const double t = CONE ? sj * tau_w : 1.0;
                acc_top += lds_interp(G + lane * kGStride, vp.beta0 + t * ztop, ZN) * qj;
            }
        }
    }
    if constexpr (kSumThenDiff) {
        // row v = boundary v+1 - boundary v; lane 63 gets 0 - boundary 63
#pragma unroll
        for (int q = 0; q < NU; ++q) acc[q] = shift_down1(acc[q]) - acc[q];
    }
    if (defer_top) {
#pragma unroll
        for (int q = 0; q < NU; ++q) {
            const T top = readlane_t(acc_top, q);
            if (lane == 63) acc[q] += top;
        }
    }

    if (lane < DNV) {
        const double ezv = g.ez[lane];
        const double rf = g.rowfac[lane];
        CTLAB_GMEM float* __restrict__ out = prj + (long)p * g.DNU * DNV + lane;
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


// inclusive v-prefix of y*L for one (view, column) row idx = p*DNU + u (lane = v)
// With real / rowSum (SART weighting, src/SARTWeighting.h:22-37, eps = 1e-7f),
// y is first replaced by (real - y) / rowSum, 0 where rowSum <= 1e-7.
template <typename T>
__device__ __forceinline__ void prjprefix_wave(const DevGeom& g, const CTLAB_GMEM float* __restrict__ prj,
                                               CTLAB_GMEM T* __restrict__ P, long idx, int lane,
                                               const CTLAB_GMEM float* __restrict__ real = nullptr,
                                               const CTLAB_GMEM float* __restrict__ rowSum = nullptr)
{
    const int DNV = g.DNV;
    T v = T(0);
    if (lane < DNV) {
        const double ezv = g.ez[lane];
        const double L = g.dx * sqrt(g.exy2[idx] + ezv * ezv) * fabs(g.tau[idx]);
        float y = prj[idx * DNV + lane];
        if (real) {
            const float w = rowSum[idx * DNV + lane];
            y = w > 1.0e-7f ? (real[idx * DNV + lane] - y) / w : 0.0f;
        }
        v = (T)((double)y * L);
    }
    const T r = wave_inclusive_scan(v);
    if (lane < DNV) P[idx * DNV + lane] = r;
}

// z-mode of the backprojector's column step
enum ZMode { kZFan = 0, kZCone = 1, kZCone64 = 2 };   // multi-slice fan; cone, ZN < 64; cone, ZN == 64

__host__ __device__ inline int zmode_of(const DevGeom& g) { return !g.cone ? kZFan : g.ZN == 64 ? kZCone64 : kZCone; }

// Adjoint z-part for one detector column: sum_v y*L * (fraction of cell v
// covered by the voxel z-interval [lane, lane+1] mapped onto the detector rows),
// read from the column's v-prefix Pu (lane v holds prefix v+1) through LDS.
// esdz = (cell-centre slab coordinate - source) / row pitch, so that the
// voxel boundary b maps to the fractional row (b - beta0) * esdz / sj - zhat0/dzh.
// Cone beam: every lane interpolates its lower voxel boundary only and takes
// the upper one from lane z+1 (DPP shift).  Lane 63 of a 64-slice volume has no
// neighbour: its upper boundary (z = 64) mostly projects off the detector,
// where the prefix is the column total (above) or 0 (below), and is
// interpolated only when it lands on the detector; the DPP shift leaves lane 63
// its own fill value.  The column total is G[DNV], which lane 63 holds as Pu
// (its v index clamps to DNV - 1).  top_total: the view's top boundary is
// above the detector for every slab and column (decided on the host).  The differences follow the
// detector row axis; the caller orients them along z (sign of the row pitch).
template <int ZM, typename T>
__device__ __forceinline__ T column_delta(const DevGeom& g, T* G, T Pu, double esdz, double lzb, double zoff,
                                          double inv_sj, double beta0, int top_total, int lane)
{
    const int DNV = g.DNV;
    G[lane + 1] = Pu;
    if constexpr (ZM == kZFan) {   // multi-slice fan beam: boundaries are the integer rows
        return lds_interp(G, (double)(lane + 1), DNV) - lds_interp(G, (double)lane, DNV);
    } else {
        const T Q = lds_interp(G, esdz * lzb - zoff, DNV);
        if constexpr (ZM == kZCone) {
            return shift_down1(Q) - Q;
        } else {
            T top;   // only lane 63's value is used
            if (top_total) {
                top = Pu;
            } else {
                const double ntop = (64.0 - beta0) * inv_sj * esdz - zoff;
                top = ntop >= (double)DNV ? Pu : ntop <= 0.0 ? T(0) : lds_interp(G, ntop, DNV);
            }
            return shift_down1_fill(Q, top) - Q;
        }
    }
}

// Column record in walk order (increasing row coordinate a): slopes of the
// column's lower and upper boundary rays times sign(s0) (so that, with
// |rs| = |slab distance| / dx, a = rbase + |rs| * slope grows with the slope),
// 1/|slope step|, and esdz.  Each view holds DNU records and a sentinel (upper
// slope NaN, 1/|slope step| and esdz 0), and one spare record follows the last
// view.
struct ColRec { double klo, khi, invdk, esdz; };

// Backprojector walk of one (slab j, tile rt of NM rows) over the views of one
// slab orientation; lane = z.  acc[i] (zeroed here) receives voxel row
// rt*NM + i.  See dd_back_walk for the structure.
// Row accumulator type in LDS: FP64 up to 12-row tiles; FP32 for 16-row tiles,
// whose FP64 accumulators (8 KB per wavefront) would cap occupancy at 4
// wavefronts per SIMD.  Each view's contribution is still formed in FP64.
template <typename T, int NM> struct BackAcc { typedef T type; };
template <> struct BackAcc<double, 16> { typedef float type; };

// A: (NM + 1) * 64 accumulators of LDS; row i lives at A[(i + 1) * 64 + lane],
// row -1 is a spare that takes the walk's first (unpaired) row boundary.
template <typename T, int NM, int ZM>
__device__ __forceinline__ void back_walk_accum(const DevGeom& g, const CTLAB_GMEM T* __restrict__ P,
                                                const CTLAB_GMEM int* __restrict__ views, int nviews, int NR, int j,
                                                int rt, const CTLAB_GMEM ColRec* __restrict__ rec,
                                                const CTLAB_GMEM unsigned short* __restrict__ start,
                                                int SJ, int SR, int lane, T* G,
                                                typename BackAcc<T, NM>::type* A, T (&acc)[NM])
{
    typedef typename BackAcc<T, NM>::type AccT;
    if (lane == 0) G[0] = T(0);
    // row accumulators live in LDS: the row index varies at run time, and
    // indexing registers with it costs four scalar-set / move pairs per FP64
    // update
    A += 64;
#pragma unroll
    for (int i = 0; i < NM; ++i) A[i * 64 + lane] = 0;
    const int r0 = rt * NM;
    const int nm = min(NM, NR - r0);
    const int DNU = g.DNU, DNV = g.DNV;
    const int vl = min(lane, DNV - 1);
    const long srow = (long)j * SR + (r0 >> 3);
    const T sgn = g.dzh < 0.0 ? T(-1) : T(1);
    const double zoff = g.zhat0 / g.dzh;

    // View pipeline: the view list is read three views ahead, the start
    // columns two ahead and the v-prefix entries of a view's first two
    // columns one ahead, so that a view's first column does not wait on
    // memory.  They are read through the view's buffer resource, so a view
    // whose start is 0xFFFF (skipped) reads zeros instead of out of bounds.
    auto first_two = [&](int pq, int sq, T& a, T& b) {
        const int asc = g.vp[pq].asc;
        const __amdgpu_buffer_rsrc_t rsq = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(P + (long)pq * DNU * DNV), (short)0, DNU * DNV * (int)sizeof(T), kBufWord3);
        const unsigned off = (unsigned)((asc ? sq : DNU - 1 - sq) * DNV + vl) * (unsigned)sizeof(T);
        a = buf_load<T>(rsq, off);
        b = buf_load<T>(rsq, off + (unsigned)((asc ? DNV : -DNV) * (int)sizeof(T)));
    };
    const long sstride = (long)SJ * SR;
    int p1 = nviews > 0 ? views[0] : 0;   // views iv, iv + 1, iv + 2 at the top of iteration iv
    int p2 = nviews > 1 ? views[1] : 0;
    int p3 = nviews > 2 ? views[2] : 0;
    int st1 = nviews > 0 ? start[p1 * sstride + srow] : 0xFFFF;
    int st2 = nviews > 1 ? start[p2 * sstride + srow] : 0xFFFF;
    T Q0 = T(0), Q1 = T(0);
    if (nviews > 0) first_two(__builtin_amdgcn_readfirstlane(p1), __builtin_amdgcn_readfirstlane(st1), Q0, Q1);
    for (int iv = 0; iv < nviews; ++iv) {
        const int p = __builtin_amdgcn_readfirstlane(p1);
        const int st = __builtin_amdgcn_readfirstlane(st1);
        const T Pr0 = Q0;
        T Pr1 = Q1;
        p1 = p2;
        st1 = st2;
        p2 = p3;
        if (iv + 1 < nviews)
            first_two(__builtin_amdgcn_readfirstlane(p1), __builtin_amdgcn_readfirstlane(st1), Q0, Q1);
        if (iv + 2 < nviews) st2 = start[p2 * sstride + srow];
        if (iv + 3 < nviews) p3 = views[iv + 3];
        if (st == 0xFFFF) continue;
        const ViewParams vp = g.vp[p];
        const double inv_sj = g.invs[(long)p * g.sjs + j];
        const double inv_rs = fabs(g.dx * inv_sj);
        const double lzb = ((double)lane - vp.beta0) * inv_sj;
        // Row boundary m in slope space: km = (m - rbase) / |rs|.  It lies at or
        // above column c's upper edge when km >= khi(c), and its fraction of the
        // column is (km - klo) / (khi - klo); both read the record's values as
        // they are, with no per-column FP64 set-up.
        double km = ((double)r0 - vp.rbase) * inv_rs;
        // walk-order column c = st: its record (its v-prefix entries and the
        // next column's came through the pipeline); each further column's are
        // loaded one column ahead.  The view's v-prefix rows are read through
        // a buffer resource, so the step past either end of the view (the
        // sentinel column) reads 0 instead of needing a clamp.
        const CTLAB_GMEM ColRec* __restrict__ rq = rec + (long)p * (DNU + 1) + st;
        const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(P + (long)p * DNU * DNV), (short)0, DNU * DNV * (int)sizeof(T), kBufWord3);
        unsigned voff = (unsigned)((vp.asc ? st : DNU - 1 - st) * DNV + vl) * (unsigned)sizeof(T);
        const unsigned vstep = (unsigned)((vp.asc ? DNV : -DNV) * (int)sizeof(T));
        ColRec r = rq[0];
        if (r.klo >= fma((double)nm, inv_rs, km)) continue;
        double klo = r.klo, khi = r.khi, iw = r.invdk;
        rq += 1;
        voff += vstep;
        ColRec rn = rq[0];
        T D = column_delta<ZM>(g, G, Pr0, r.esdz, lzb, zoff, inv_sj, vp.beta0, vp.top_total, lane);
        T Tsum = T(0), Wprev = T(0);
        // W at row boundary r0 + i: advance while the column ends at or below
        // it; the sentinel's NaN upper edge stops the walk after the last
        // column, whose term has then moved into Tsum while the sentinel adds
        // 0 (its 1/|slope step| is 0)
        auto boundary = [&]() {
            while (km >= khi) {
                Tsum += D;
                klo = khi;
                khi = rn.khi;
                iw = rn.invdk;
                D = column_delta<ZM>(g, G, Pr1, rn.esdz, lzb, zoff, inv_sj, vp.beta0, vp.top_total, lane);
                rq += 1;
                voff += vstep;
                rn = rq[0];
                Pr1 = buf_load<T>(prs, voff);
            }
            const double fr = fmin(fmax((km - klo) * iw, 0.0), 1.0);
            km += inv_rs;
            return Tsum + (T)fr * D;
        };
        // two row boundaries per trip (each copy carries its own column walk):
        // the second row's accumulator is an immediate offset from the first's,
        // and W and Wprev alternate registers
        for (int i = 0; i <= nm; i += 2) {
            AccT* Ai = A + (i - 1) * 64 + lane;
            const T W0 = boundary();
            Ai[0] += (AccT)(W0 - Wprev);
            if (i + 1 > nm) break;
            const T W1 = boundary();
            Ai[64] += (AccT)(W1 - W0);
            Wprev = W1;
        }
    }
    // the z-differences above are taken along the detector row axis; sgn
    // orients them along z
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] = (T)A[i * 64 + lane] * sgn;
}

// Writes one backprojected tile (lane = z): for voxel row i of the tile,
// store(idx, v) receives the volume index and the masked value.
template <typename T, int NM, class Store>
__device__ __forceinline__ void back_store_tile(const DevGeom& g, const T (&acc)[NM], int yslab, int NR, int j,
                                                int rt, int lane, Store&& store)
{
    const int r0 = rt * NM;
    const int nm = min(NM, NR - r0);
    if (lane < g.ZN) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            if (i >= nm) break;
            const int m = r0 + i;
            const long col = yslab ? ((long)j * g.XN + m) : ((long)m * g.XN + j);
            const float mk = (g.mask == nullptr || g.mask[col]) ? 1.0f : 0.0f;
            store(col * g.ZN + lane, (float)acc[i] * mk);
        }
    }
}

// Backprojection of one (slab j, row tile rt) over the listed views of one
// slab orientation; overwrites (accumulate = 0) or adds to vol.
template <typename T, int NM, int ZM>
__device__ __forceinline__ void back_walk_wave(const DevGeom& g, const CTLAB_GMEM T* __restrict__ P,
                                               CTLAB_GMEM float* __restrict__ vol, const CTLAB_GMEM int* __restrict__ views,
                                               int nviews, int yslab, int accumulate, int NR, int j, int rt,
                                               const CTLAB_GMEM ColRec* __restrict__ rec,
                                               const CTLAB_GMEM unsigned short* __restrict__ start, int SJ, int SR, int lane,
                                               T* G, typename BackAcc<T, NM>::type* A)
{
    T acc[NM];
    back_walk_accum<T, NM, ZM>(g, P, views, nviews, NR, j, rt, rec, start, SJ, SR, lane, G, A, acc);
    back_store_tile<T, NM>(g, acc, yslab, NR, j, rt, lane, [&](long idx, float v) {
        vol[idx] = accumulate ? vol[idx] + v : v;
    });
}

}  // namespace ctlab

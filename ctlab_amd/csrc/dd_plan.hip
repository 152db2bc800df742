// Host side: geometry plans (per-view FP64 tables on the device), the
// device-resident plan API, and CTLAB's host-pointer entry points
// DD3Proj_gpu / DD3Back_gpu / DD2_multiGPU as thin H2D -> plan -> D2H wrappers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <list>
#include <memory>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctlab_mi355x.h"
#include "dd_common.hpp"
#include "plan.hpp"
#include "status.hpp"

namespace ctlab {


// ---------------------------------------------------------------------------
// status (per host thread)
// ---------------------------------------------------------------------------
static thread_local int t_err = CTLAB_OK;
static thread_local char t_msg[512] = "";

void set_error(int code, const char* fmt, ...)
{
    t_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
}

void clear_error()
{
    t_err = CTLAB_OK;
    t_msg[0] = 0;
}

// A failed call is recorded here and taken off the runtime's per-thread last
// error, which successful calls leave in place: a later launch check
// (hipGetLastError, here or in the caller's own code) would otherwise report
// this failure as its own.
int hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return 0;
    set_error(e == hipErrorOutOfMemory ? CTLAB_ERR_NOMEM : CTLAB_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
    (void)hipGetLastError();
    return -1;
}

// CTLAB's entry points return void, and its callers never read a status: a
// failed call of one of them also prints its message to stderr (once, from the
// outermost entry point of a nested call; CTLAB_QUIET=1 in the environment
// turns it off).
static thread_local int t_void_depth = 0;
struct VoidReport {
    const char* fn;
    explicit VoidReport(const char* f) : fn(f) { ++t_void_depth; }
    ~VoidReport()
    {
        if (--t_void_depth == 0 && t_err != CTLAB_OK) {
            const char* q = getenv("CTLAB_QUIET");
            if (!(q && q[0] == '1')) fprintf(stderr, "%s: error %d: %s\n", fn, t_err, t_msg);
        }
    }
};

// src/utilities.cuh:413-433 (DD3Boundaries)
static void boundaries(int nb, const double* c, double* b)
{
    if (nb >= 3) {
        b[0] = 1.5 * c[0] - 0.5 * c[1];
        for (int i = 1; i <= nb - 2; ++i) b[i] = 0.5 * c[i - 1] + 0.5 * c[i];
        b[nb - 1] = 1.5 * c[nb - 2] - 0.5 * c[nb - 3];
    } else {
        b[0] = c[0] - 0.5;
        b[1] = c[0] + 0.5;
    }
}

// src/utilities.cuh:279-286 (regularizeAngle)
static double regularize_angle(double a)
{
    const double twopi = 6.283185307179586;
    while (a >= twopi) a -= twopi;
    while (a < 0) a += twopi;
    return a;
}

}  // namespace ctlab

namespace ctlab {

static bool finite_all(const std::vector<double>& v)
{
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// FP64 per-view tables; same formulas as the CPU oracle (dd_oracle.c).
static int build_tables(ctlab_plan* P)
{
    const GeomIn& g = P->in;
    if (g.DNU <= 0 || g.DNV <= 0 || g.PN <= 0 || g.XN <= 0 || g.YN <= 0 || g.ZN <= 0) {
        set_error(CTLAB_ERR_ARG, "sizes must be positive (DNU=%d DNV=%d PN=%d XN=%d YN=%d ZN=%d)",
                  g.DNU, g.DNV, g.PN, g.XN, g.YN, g.ZN);
        return -1;
    }
    if (!(g.dx > 0) || !(g.dz > 0) || !std::isfinite(g.dx) || !std::isfinite(g.dz)) {
        set_error(CTLAB_ERR_ARG, "voxel pitches must be positive and finite");
        return -1;
    }
    if (!std::isfinite(g.x0) || !std::isfinite(g.y0) || !std::isfinite(g.z0) || !std::isfinite(g.cx) ||
        !std::isfinite(g.cy) || !std::isfinite(g.cz) || !finite_all(g.xds) || !finite_all(g.yds) ||
        !finite_all(g.zds) || !finite_all(g.angs) || !finite_all(g.zpos)) {
        set_error(CTLAB_ERR_ARG, "non-finite geometry");
        return -1;
    }
    if (!g.cone && g.ZN != g.DNV) {
        set_error(CTLAB_ERR_ARG, "multi-slice mode needs ZN == DNV");
        return -1;
    }
    const int DNU = g.DNU, DNV = g.DNV, PN = g.PN;
    std::vector<double> xb(DNU + 1), yb(DNU + 1), zb(DNV + 1);
    boundaries(DNU + 1, g.xds.data(), xb.data());
    boundaries(DNU + 1, g.yds.data(), yb.data());
    if (g.cone) {
        if (g.zbnd.size() == (size_t)DNV + 1) zb = g.zbnd;
        else boundaries(DNV + 1, g.zds.data(), zb.data());
    }

    P->zhat.resize(DNV + 1);
    P->ez.resize(DNV);
    P->rowfac.resize(DNV);
    for (int v = 0; v <= DNV; ++v) P->zhat[v] = g.cone ? (zb[v] - g.z0) / g.dz : (double)v;
    for (int v = 0; v < DNV; ++v) {
        P->ez[v] = g.cone ? (g.zds[v] - g.z0) : 0.0;
        const double h = P->zhat[v + 1] - P->zhat[v];
        if (h == 0.0) {
            set_error(CTLAB_ERR_GEOMETRY, "zero-height detector row %d", v);
            return -1;
        }
        P->rowfac[v] = g.cone ? 1.0 / h : 1.0;
    }
    // uniform row model for the backprojector (rows must be evenly spaced)
    P->zhat0 = P->zhat[0];
    P->dzh = (P->zhat[DNV] - P->zhat[0]) / DNV;
    for (int v = 0; v <= DNV; ++v) {
        const double dev = std::fabs(P->zhat[v] - (P->zhat0 + v * P->dzh));
        if (dev > 1e-4 * std::fabs(P->dzh)) {
            set_error(CTLAB_ERR_GEOMETRY, "detector rows are not evenly spaced (row %d off by %g rows)",
                      v, dev / std::fabs(P->dzh));
            return -1;
        }
    }

    const double Xmin = g.cx - g.XN * g.dx * 0.5;
    const double Ymin = g.cy - g.YN * g.dx * 0.5;
    const double Zmin = g.cz - g.ZN * g.dz * 0.5;
    P->vp.resize(PN);
    P->kap.resize((size_t)PN * (DNU + 1));
    P->tau.resize((size_t)PN * DNU);
    P->colfac.resize((size_t)PN * DNU);
    P->exy2.resize((size_t)PN * DNU);
    P->invdk.resize((size_t)PN * DNU);
    P->esdz.resize((size_t)PN * DNU);
    P->invs.assign((size_t)PN * std::max(g.XN, g.YN), 0.0);
    // DNU walk-order records and a sentinel per view, one spare record after the last view
    P->rec.assign(((size_t)PN * (DNU + 1) + 1) * 4, 0.0);
    if (DNU >= 0xFFFF) {
        set_error(CTLAB_ERR_GEOMETRY, "DNU (%d) must be below 65535", DNU);
        return -1;
    }
    P->views_y.clear();
    P->views_x.clear();
    for (int p = 0; p < PN; ++p) {
        const double th = regularize_angle(g.angs[p]);
        const double c = std::cos(th), s = std::sin(th);
        const double Sx = g.x0 * c - g.y0 * s, Sy = g.x0 * s + g.y0 * c;
        const double zs = g.z0 + g.zpos[p];
        const int yslab = std::fabs(g.cy - Sy) >= std::fabs(g.cx - Sx);
        ViewParams& V = P->vp[p];
        V.yslab = yslab;
        double Ss, Sr, Smin, Rmin;
        if (yslab) {
            V.NS = g.YN; V.NR = g.XN; V.slab_stride = g.XN; V.row_stride = 1;
            Ss = Sy; Sr = Sx; Smin = Ymin; Rmin = Xmin;
        } else {
            V.NS = g.XN; V.NR = g.YN; V.slab_stride = 1; V.row_stride = g.XN;
            Ss = Sx; Sr = Sy; Smin = Xmin; Rmin = Ymin;
        }
        V.s0 = Smin + 0.5 * g.dx - Ss;
        V.rbase = (Sr - Rmin) / g.dx;
        V.beta0 = g.cone ? (zs - Zmin) / g.dz : 0.0;
        const double slast = V.s0 + (V.NS - 1) * g.dx;
        if (!(V.s0 * slast > 0)) {
            set_error(CTLAB_ERR_GEOMETRY, "view %d: source lies inside the volume's slab range", p);
            return -1;
        }
        double* kp = &P->kap[(size_t)p * (DNU + 1)];
        for (int u = 0; u <= DNU; ++u) {
            const double X = xb[u] * c - yb[u] * s, Y = xb[u] * s + yb[u] * c;
            const double ds = yslab ? Y : X, dr = yslab ? X : Y;
            kp[u] = (dr - Sr) / (ds - Ss);
        }
        for (int u = 0; u <= DNU; ++u) {
            if (!std::isfinite(kp[u])) {
                set_error(CTLAB_ERR_GEOMETRY, "view %d: detector boundary %d is level with the source", p, u);
                return -1;
            }
        }
        const double dk0 = kp[1] - kp[0];
        for (int u = 0; u < DNU; ++u) {
            const double dk = kp[u + 1] - kp[u];
            if (!(dk * dk0 > 0)) {
                set_error(CTLAB_ERR_GEOMETRY, "view %d: detector boundaries are not monotone", p);
                return -1;
            }
            const double X = g.xds[u] * c - g.yds[u] * s, Y = g.xds[u] * s + g.yds[u] * c;
            const double ex = X - Sx, ey = Y - Sy;
            const double es = yslab ? ey : ex;
            if (es == 0.0) {
                set_error(CTLAB_ERR_GEOMETRY, "view %d: detector cell %d is level with the source", p, u);
                return -1;
            }
            const double t = 1.0 / es;
            P->tau[(size_t)p * DNU + u] = t;
            P->exy2[(size_t)p * DNU + u] = ex * ex + ey * ey;
            P->invdk[(size_t)p * DNU + u] = 1.0 / std::fabs(dk);
            P->esdz[(size_t)p * DNU + u] = es / P->dzh;
            P->colfac[(size_t)p * DNU + u] =
                g.cone ? g.dx * g.dx / std::fabs(dk) : g.dx * g.dx * std::fabs(t) / std::fabs(dk);
        }
        V.asc = (V.s0 * dk0) > 0;
        const int sjs = std::max(g.XN, g.YN);
        for (int j = 0; j < V.NS; ++j) P->invs[(size_t)p * sjs + j] = 1.0 / (V.s0 + j * g.dx);
        // row coordinate of the top voxel boundary, (ZN - beta0) * esdz / sj - zhat0 / dzh, is
        // bilinear in (esdz, 1/sj): its corners bound it over all slabs and columns
        V.top_total = 0;
        if (g.cone) {
            double emin = P->esdz[(size_t)p * DNU], emax = emin;
            for (int u = 1; u < DNU; ++u) {
                emin = std::min(emin, P->esdz[(size_t)p * DNU + u]);
                emax = std::max(emax, P->esdz[(size_t)p * DNU + u]);
            }
            double lo = HUGE_VAL;
            for (double e : {emin, emax})
                for (double sj : {V.s0, slast})
                    lo = std::min(lo, (g.ZN - V.beta0) * e / sj - P->zhat0 / P->dzh);
            V.top_total = lo >= (double)DNV;
        }
        (yslab ? P->views_y : P->views_x).push_back(p);
        // walk-order (increasing row coordinate) column records for the backprojector.
        // Slopes are stored times the sign of the source-to-slab distance, which is
        // the same for every slab of the view: a boundary with stored slope k' lies
        // at row coordinate a = rbase + |rs| * k', so k' grows in walk order and
        // the walk compares row boundaries with column edges in slope space.
        const double sgn_s = V.s0 > 0 ? 1.0 : -1.0;
        for (int c = 0; c < DNU; ++c) {
            const int u = V.asc ? c : DNU - 1 - c;
            double* r = &P->rec[((size_t)p * (DNU + 1) + c) * 4];
            r[0] = sgn_s * (V.asc ? kp[u] : kp[u + 1]);
            r[1] = sgn_s * (V.asc ? kp[u + 1] : kp[u]);
            r[2] = P->invdk[(size_t)p * DNU + u];
            r[3] = P->esdz[(size_t)p * DNU + u];
        }
        // sentinel: NaN slopes stop the column walk (no row boundary compares
        // at or above NaN); zero 1/|slope step| and esdz keep its term finite
        double* rs = &P->rec[((size_t)p * (DNU + 1) + DNU) * 4];
        rs[0] = rs[1] = std::numeric_limits<double>::quiet_NaN();
        rs[2] = rs[3] = 0.0;
    }
    return 0;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src)
{
    if (src.empty()) { *dst = nullptr; return 0; }
    if (hip_check(hipMalloc((void**)dst, src.size() * sizeof(T)), "hipMalloc")) return -1;
    return hip_check(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice),
                     "hipMemcpy");
}

void free_plan(ctlab_plan* P)
{
    if (!P) return;
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(P->device);
    for (auto& pt : P->parts) free_plan(pt.plan);
    P->parts.clear();
    if (P->sub_shared) P->d_sub_vol = P->d_sub_prj = nullptr;
    if (P->d_sub_vol) hipFree(P->d_sub_vol);
    if (P->d_sub_prj) hipFree(P->d_sub_prj);
    if (P->ws_shared) P->d_cz = P->d_P = nullptr;
    void* ptrs[] = {P->d_vp, P->d_kap, P->d_tau, P->d_colfac, P->d_exy2, P->d_invdk, P->d_esdz, P->d_invs, P->d_zhat, P->d_ez, P->d_rec, P->d_start,
                    P->d_start1, P->d_rowfac, P->d_mask, P->d_views, P->d_cz, P->d_P};
    for (void* q : ptrs)
        if (q) hipFree(q);
    hipSetDevice(cur);
    delete P;
}

// Workspace elements a plan needs: its integral image and projection v-prefix,
// or, for a blocked plan, those of its largest part (the parts run one after
// another on one workspace).
static void ws_need(const GeomIn& in, size_t& ncz, size_t& nP)
{
    GeomIn g;
    g.XN = in.XN; g.YN = in.YN;
    g.ZN = std::min(in.ZN, kBlock);
    g.DNV = std::min(in.DNV, kBlock);
    g.cone = in.cone;
    ncz = cz_elems(g);
    nP = (size_t)in.PN * in.DNU * std::min(in.DNV, kBlock);
    if (is_fan2d(in)) nP = std::max(nP, (size_t)in.ZN * in.PN * (in.DNU + 1));   // the columns' prefix
}

// (Re)allocates the integral-image and v-prefix workspaces for `bits`.  Both
// new buffers are allocated before the old ones are released, so on failure
// the plan keeps its previous workspace and precision unchanged.
static int alloc_workspace(ctlab_plan* P, int bits)
{
    const size_t es = bits == 32 ? 4 : 8;
    size_t ncz = 0, nP = 0;
    ws_need(P->in, ncz, nP);
    void *cz = nullptr, *pp = nullptr;
    if (hip_check(hipMalloc(&cz, ncz * es), "hipMalloc(integral image)")) return -1;
    if (hip_check(hipMalloc(&pp, nP * es), "hipMalloc(projection prefix)")) {
        hipFree(cz);
        return -1;
    }
    if (!P->ws_shared) {
        if (P->d_cz) hipFree(P->d_cz);
        if (P->d_P) hipFree(P->d_P);
    }
    P->d_cz = cz;
    P->d_P = pp;
    P->ws_shared = false;
    P->sat_bits = bits;
    P->ws_elem = es;
    P->ws_cz_cap = ncz;
    P->ws_P_cap = nP;
    return 0;
}

// W's workspace can serve P: same device and precision, large enough
static bool can_share(const ctlab_plan* P, const ctlab_plan* W)
{
    if (!W || !W->d_cz || !W->d_P || W->device != P->device || W->sat_bits != P->sat_bits) return false;
    size_t ncz = 0, nP = 0;
    ws_need(P->in, ncz, nP);
    return W->ws_cz_cap >= ncz && W->ws_P_cap >= nP;
}

static void borrow_workspace(ctlab_plan* P, const ctlab_plan* W)
{
    P->d_cz = W->d_cz;
    P->d_P = W->d_P;
    P->ws_elem = W->ws_elem;
    P->ws_cz_cap = W->ws_cz_cap;
    P->ws_P_cap = W->ws_P_cap;
    P->ws_shared = true;
}

static ctlab_plan* finish_blocked(ctlab_plan* P, int gpunum, const ctlab_plan* share_ws);

ctlab_plan* finish_plan(ctlab_plan* P, int gpunum, const ctlab_plan* share_ws)
{
    if (is_blocked(P->in)) return finish_blocked(P, gpunum, share_ws);
    if (build_tables(P)) { delete P; return nullptr; }
    {
        // A launch holds at most 2^32 work-items (256 per block here).  The
        // walks have one block per 4 views and tile of columns (projector) and
        // per 4 slabs and tile of rows (backprojector); at the smallest tiles
        // (8 columns, 8 rows) their grids are the largest any plan launches.
        const GeomIn& in = P->in;
        const long lim = 0xFFFFFFFFL / 256;
        const long g8 = ((in.PN + 3L) / 4 + 7) / 8;
        const long proj_blocks = 8 * g8 * ((in.DNU + 7L) / 8);
        const long back_blocks = ((in.XN + 7L) / 8) * ((in.YN + 3L) / 4) + ((in.YN + 7L) / 8) * ((in.XN + 3L) / 4);
        if (proj_blocks > lim || back_blocks > lim) {
            set_error(CTLAB_ERR_GEOMETRY,
                      "%ld projector / %ld backprojector blocks at 8-column and 8-row tiles: more than one launch "
                      "holds (%ld); fewer views x columns or a smaller XN x YN",
                      proj_blocks, back_blocks, lim);
            delete P;
            return nullptr;
        }
    }
    int ndev = 0;
    if (hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount")) { delete P; return nullptr; }
    if (gpunum < 0 || gpunum >= ndev) {
        set_error(CTLAB_ERR_ARG, "gpunum %d out of range (%d devices)", gpunum, ndev);
        delete P;
        return nullptr;
    }
    P->device = gpunum;
    if (hip_check(hipSetDevice(gpunum), "hipSetDevice")) { delete P; return nullptr; }
    std::vector<int> views(P->views_y);
    views.insert(views.end(), P->views_x.begin(), P->views_x.end());
    if (upload(&P->d_vp, P->vp) || upload(&P->d_kap, P->kap) || upload(&P->d_tau, P->tau) ||
        upload(&P->d_colfac, P->colfac) || upload(&P->d_exy2, P->exy2) || upload(&P->d_invdk, P->invdk) ||
        upload(&P->d_esdz, P->esdz) || upload(&P->d_zhat, P->zhat) || upload(&P->d_invs, P->invs) ||
        upload(&P->d_ez, P->ez) || upload(&P->d_rowfac, P->rowfac) || upload(&P->d_mask, P->in.mask) ||
        upload(&P->d_views, views) || upload(&P->d_rec, P->rec) ||
        (can_share(P, share_ws) ? (borrow_workspace(P, share_ws), 0) : alloc_workspace(P, P->sat_bits))) {
        free_plan(P);
        return nullptr;
    }
    // backprojector start table (per view, slab, 8-row group), built once on the device
    P->SJ = std::max(P->in.XN, P->in.YN);
    P->SR = (P->SJ + 7) / 8 + 1;
    const size_t nst = (size_t)P->in.PN * P->SJ * P->SR;
    if (hip_check(hipMalloc((void**)&P->d_start, nst * sizeof(unsigned short)), "hipMalloc(start table)") ||
        hip_check(launch_back_start(P->dev(), P->d_rec, P->d_start, P->SJ, P->SR, nullptr), "start table") ||
        hip_check(hipDeviceSynchronize(), "start table")) {
        free_plan(P);
        return nullptr;
    }
    // fan beam of up to kFanSlices slices: the one-row start table of its
    // backprojector, up to 1 GiB (above that the backprojector takes the
    // general walk)
    const size_t ns1 = (size_t)P->in.PN * P->SJ * (P->SJ + 1);
    if (is_fan2d(P->in) && ns1 * sizeof(unsigned short) <= ((size_t)1 << 30)) {
        P->SR1 = P->SJ + 1;
        if (hip_check(hipMalloc((void**)&P->d_start1, ns1 * sizeof(unsigned short)), "hipMalloc(row start table)") ||
            hip_check(launch_fan2d_start(P->dev(), P->d_rec, P->d_start1, P->SJ, P->SR1, nullptr), "row start table") ||
            hip_check(hipDeviceSynchronize(), "row start table")) {
            free_plan(P);
            return nullptr;
        }
    }
    P->rec.clear();
    P->rec.shrink_to_fit();
    return P;
}

// ---------------------------------------------------------------------------
// Blocked plans: ZN > 64 or DNV > 64.  The projection of slice block c onto
// row block b uses only those slices and rows, so A = sum over (c, b) of the
// block operators A_cb, each a plan of at most 64 slices and 64 rows: its
// volume is the slice block in place (image centre moved to the block's
// centre), its detector the row block with the row boundaries of the whole
// detector.  Pairs whose rays cannot meet are left out (their A_cb is 0: every
// row boundary maps below the slice block or above it, where the row
// differences cancel).  Multi-slice fan beam: slice block c meets row block c.
// ---------------------------------------------------------------------------

// true if, in some view, a ray through row block [v0, v0+nv] can cross the
// slice block [z0, z0+nz] (lattice units; conservative by one slice)
static bool pair_meets(const GeomIn& g, const std::vector<double>& zhat, int z0, int nz, int v0, int nv)
{
    const double Xmin = g.cx - g.XN * g.dx * 0.5, Ymin = g.cy - g.YN * g.dx * 0.5;
    const double Zmin = g.cz - g.ZN * g.dz * 0.5;
    const double za = zhat[v0], zb = zhat[v0 + nv];
    for (int p = 0; p < g.PN; ++p) {
        const double th = regularize_angle(g.angs[p]);
        const double c = std::cos(th), s = std::sin(th);
        const double Sx = g.x0 * c - g.y0 * s, Sy = g.x0 * s + g.y0 * c;
        const int yslab = std::fabs(g.cy - Sy) >= std::fabs(g.cx - Sx);
        const int NS = yslab ? g.YN : g.XN;
        const double s0 = (yslab ? Ymin - Sy : Xmin - Sx) + 0.5 * g.dx, s1 = s0 + (NS - 1) * g.dx;
        double emin = HUGE_VAL, emax = -HUGE_VAL;
        for (int u = 0; u < g.DNU; ++u) {
            const double X = g.xds[u] * c - g.yds[u] * s, Y = g.xds[u] * s + g.yds[u] * c;
            const double es = yslab ? Y - Sy : X - Sx;
            emin = std::min(emin, es);
            emax = std::max(emax, es);
        }
        if (!(emin * emax > 0) || !(s0 * s1 > 0)) return true;   // no sign-definite bound: keep
        // z magnification t = s / e over slabs and columns, then z = beta0 + t * zhat
        double tlo = HUGE_VAL, thi = -HUGE_VAL;
        for (double sv : {s0, s1})
            for (double ev : {emin, emax}) {
                tlo = std::min(tlo, sv / ev);
                thi = std::max(thi, sv / ev);
            }
        const double beta0 = (g.z0 + g.zpos[p] - Zmin) / g.dz;
        const double zlo = beta0 + std::min(std::min(tlo * za, tlo * zb), std::min(thi * za, thi * zb));
        const double zhi = beta0 + std::max(std::max(tlo * za, tlo * zb), std::max(thi * za, thi * zb));
        if (zhi > z0 - 1.0 && zlo < z0 + nz + 1.0) return true;
    }
    return false;
}

static ctlab_plan* finish_blocked(ctlab_plan* P, int gpunum, const ctlab_plan* share_ws)
{
    const GeomIn& g = P->in;
    if (g.DNU <= 0 || g.DNV <= 0 || g.PN <= 0 || g.XN <= 0 || g.YN <= 0 || g.ZN <= 0 || !(g.dx > 0) ||
        !(g.dz > 0)) {
        set_error(CTLAB_ERR_ARG, "sizes and pitches must be positive");
        delete P;
        return nullptr;
    }
    if (!g.cone && g.ZN != g.DNV) {
        set_error(CTLAB_ERR_ARG, "multi-slice mode needs ZN == DNV");
        delete P;
        return nullptr;
    }
    int ndev = 0;
    if (hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount")) { delete P; return nullptr; }
    if (gpunum < 0 || gpunum >= ndev) {
        set_error(CTLAB_ERR_ARG, "gpunum %d out of range (%d devices)", gpunum, ndev);
        delete P;
        return nullptr;
    }
    P->device = gpunum;
    if (hip_check(hipSetDevice(gpunum), "hipSetDevice")) { delete P; return nullptr; }
    const int DNV = g.DNV;
    std::vector<double> zb(DNV + 1), zhat(DNV + 1);
    if (g.cone) {
        if (g.zbnd.size() == (size_t)DNV + 1) zb = g.zbnd;
        else boundaries(DNV + 1, g.zds.data(), zb.data());
        for (int v = 0; v <= DNV; ++v) zhat[v] = (zb[v] - g.z0) / g.dz;
    }
    // one workspace for every part, and the staging buffers
    bool ok = true;
    if (can_share(P, share_ws)) borrow_workspace(P, share_ws);
    else ok = !alloc_workspace(P, P->sat_bits);
    // a slice block's backprojection, a row block's projection (both are also
    // where parts that add into an output block put their share)
    const size_t sub_vol = (size_t)g.XN * g.YN * std::min(g.ZN, kBlock);
    const size_t sub_prj = (size_t)g.PN * g.DNU * std::min(DNV, kBlock);
    if (ok && share_ws && share_ws->d_sub_vol && share_ws->d_sub_prj && share_ws->device == P->device &&
        (size_t)share_ws->in.XN * share_ws->in.YN >= (size_t)g.XN * g.YN &&
        (size_t)share_ws->in.PN * share_ws->in.DNU >= (size_t)g.PN * g.DNU &&
        std::min(share_ws->in.ZN, kBlock) >= std::min(g.ZN, kBlock) &&
        std::min(share_ws->in.DNV, kBlock) >= std::min(DNV, kBlock)) {
        P->d_sub_vol = share_ws->d_sub_vol;
        P->d_sub_prj = share_ws->d_sub_prj;
        P->sub_shared = true;
    } else if (ok) {
        ok = !hip_check(hipMalloc((void**)&P->d_sub_vol, sub_vol * sizeof(float)), "hipMalloc(slice block)") &&
             !hip_check(hipMalloc((void**)&P->d_sub_prj, sub_prj * sizeof(float)), "hipMalloc(row block)");
    }
    const double Zmin = g.cz - g.ZN * g.dz * 0.5;
    const int nzb = (g.ZN + kBlock - 1) / kBlock, nvb = (DNV + kBlock - 1) / kBlock;
    for (int c = 0; ok && c < nzb; ++c) {
        const int z0 = c * kBlock, nz = std::min(kBlock, g.ZN - z0);
        for (int b = 0; ok && b < nvb; ++b) {
            const int v0 = b * kBlock, nv = std::min(kBlock, DNV - v0);
            if (g.cone ? !pair_meets(g, zhat, z0, nz, v0, nv) : c != b) continue;
            ctlab_plan* S = new ctlab_plan();
            S->in = g;
            S->in.ZN = nz;
            S->in.DNV = nv;
            S->in.zds.assign(g.zds.begin() + v0, g.zds.begin() + v0 + nv);
            if (g.cone) {
                S->in.cz = Zmin + (z0 + 0.5 * nz) * g.dz;
                S->in.zbnd.assign(zb.begin() + v0, zb.begin() + v0 + nv + 1);
            }
            S->sat_bits = P->sat_bits;
            S->tile_nu = P->tile_nu;
            S->tile_nm = P->tile_nm;
            S = finish_plan(S, gpunum, P);
            if (!S) { ok = false; break; }
            P->parts.push_back({S, c, b, z0, nz, v0, nv});
        }
    }
    if (!ok) {
        free_plan(P);
        return nullptr;
    }
    if (!P->parts.empty()) {
        P->views_y = P->parts[0].plan->views_y;
        P->views_x = P->parts[0].plan->views_x;
    }
    return P;
}

static void sync_part(const ctlab_plan* P, ctlab_plan* S)
{
    S->tile_nu = P->tile_nu;
    S->tile_nm = P->tile_nm;
}

// projection of a blocked plan: slice block by slice block (one gather each),
// each part's rows written into its row block (the first part of a row block
// writes, later ones add); row blocks no part reaches are zeroed
static int blocked_proj(ctlab_plan* P, const float* d_vol, float* d_prj, hipStream_t s)
{
    const GeomIn& in = P->in;
    const long ncol = (long)in.XN * in.YN, nrow = (long)in.PN * in.DNU;
    const int nvb = (in.DNV + kBlock - 1) / kBlock;
    std::vector<char> done(nvb, 0);
    int cur = -1;
    for (auto& pt : P->parts) {
        sync_part(P, pt.plan);
        const float* src = d_vol;
        if (pt.nz != in.ZN) {
            if (pt.zb != cur && hip_check(launch_rows_gather(d_vol, P->d_sub_vol, ncol, in.ZN, pt.z0, pt.nz, s),
                                          "slice block gather"))
                return t_err;
            cur = pt.zb;
            src = P->d_sub_vol;
        }
        const bool first = !done[pt.vb];
        if (first && pt.nv == in.DNV) {
            if (plan_proj(pt.plan, src, d_prj, s)) return t_err;
        } else {
            if (plan_proj(pt.plan, src, P->d_sub_prj, s) ||
                hip_check(launch_rows_scatter(P->d_sub_prj, d_prj, nrow, pt.nv, in.DNV, pt.v0, !first, s),
                          "row block scatter"))
                return t_err;
        }
        done[pt.vb] = 1;
    }
    for (int b = 0; b < nvb; ++b)
        if (!done[b] && hip_check(launch_rows_scatter(nullptr, d_prj, nrow, std::min(kBlock, in.DNV - b * kBlock),
                                                      in.DNV, b * kBlock, 0, s), "row block clear"))
            return t_err;
    return CTLAB_OK;
}

// backprojection of a blocked plan: row block by row block (one gather each),
// each part's slices written into its slice block (first writes, later add)
static int blocked_back(ctlab_plan* P, const float* d_prj, float* d_vol, hipStream_t s)
{
    const GeomIn& in = P->in;
    const long ncol = (long)in.XN * in.YN, nrow = (long)in.PN * in.DNU;
    const int nzb = (in.ZN + kBlock - 1) / kBlock, nvb = (in.DNV + kBlock - 1) / kBlock;
    std::vector<char> done(nzb, 0);
    for (int b = 0; b < nvb; ++b) {
        bool gathered = false;
        for (auto& pt : P->parts) {
            if (pt.vb != b) continue;
            sync_part(P, pt.plan);
            const float* src = d_prj;
            if (pt.nv != in.DNV) {
                if (!gathered && hip_check(launch_rows_gather(d_prj, P->d_sub_prj, nrow, in.DNV, pt.v0, pt.nv, s),
                                           "row block gather"))
                    return t_err;
                gathered = true;
                src = P->d_sub_prj;
            }
            const bool first = !done[pt.zb];
            if (first && pt.nz == in.ZN) {
                if (plan_back(pt.plan, src, d_vol, s)) return t_err;
            } else {
                if (plan_back(pt.plan, src, P->d_sub_vol, s) ||
                    hip_check(launch_rows_scatter(P->d_sub_vol, d_vol, ncol, pt.nz, in.ZN, pt.z0, !first, s),
                              "slice block scatter"))
                    return t_err;
            }
            done[pt.zb] = 1;
        }
    }
    for (int c = 0; c < nzb; ++c)
        if (!done[c] && hip_check(launch_rows_scatter(nullptr, d_vol, ncol, std::min(kBlock, in.ZN - c * kBlock), in.ZN,
                                                      c * kBlock, 0, s), "slice block clear"))
            return t_err;
    return CTLAB_OK;
}

static std::vector<double> to_d(const float* p, int n)
{
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i) v[i] = p[i];
    return v;
}

ctlab_plan* plan_create_cone(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                             const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                             float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                             int YN, int ZN, float dx, float dz, const byte* mask, int gpunum)
{
    if (!xds || !yds || !zds || !hangs || !hzPos) {
        set_error(CTLAB_ERR_ARG, "NULL geometry array");
        return nullptr;
    }
    if (DNU <= 0 || DNV <= 0 || PN <= 0 || XN <= 0 || YN <= 0 || ZN <= 0) {
        set_error(CTLAB_ERR_ARG, "sizes must be positive");
        return nullptr;
    }
    ctlab_plan* P = new ctlab_plan();
    GeomIn& g = P->in;
    g.x0 = x0; g.y0 = y0; g.z0 = z0; g.DNU = DNU; g.DNV = DNV;
    g.xds = to_d(xds, DNU); g.yds = to_d(yds, DNU); g.zds = to_d(zds, DNV);
    g.angs = to_d(hangs, PN); g.zpos = to_d(hzPos, PN);
    g.cx = imgXCenter; g.cy = imgYCenter; g.cz = imgZCenter;
    g.PN = PN; g.XN = XN; g.YN = YN; g.ZN = ZN; g.dx = dx; g.dz = dz; g.cone = 1;
    if (mask) g.mask.assign(mask, mask + (size_t)XN * YN);
    return finish_plan(P, gpunum);
}

ctlab_plan* plan_create_fan(float x0, float y0, const float* xds, const float* yds, int DNU, int SLN,
                            float imgXCenter, float imgYCenter, int XN, int YN, float dx,
                            const float* hangs, int PN, const byte* mask, int gpunum)
{
    if (!xds || !yds || !hangs) {
        set_error(CTLAB_ERR_ARG, "NULL geometry array");
        return nullptr;
    }
    if (DNU <= 0 || SLN <= 0 || PN <= 0 || XN <= 0 || YN <= 0) {
        set_error(CTLAB_ERR_ARG, "sizes must be positive");
        return nullptr;
    }
    ctlab_plan* P = new ctlab_plan();
    GeomIn& g = P->in;
    g.x0 = x0; g.y0 = y0; g.z0 = 0; g.DNU = DNU; g.DNV = SLN;
    g.xds = to_d(xds, DNU); g.yds = to_d(yds, DNU); g.zds.assign(SLN, 0.0);
    g.angs = to_d(hangs, PN); g.zpos.assign(PN, 0.0);
    g.cx = imgXCenter; g.cy = imgYCenter; g.cz = 0;
    g.PN = PN; g.XN = XN; g.YN = YN; g.ZN = SLN; g.dx = dx; g.dz = 1.0; g.cone = 0;
    if (mask) g.mask.assign(mask, mask + (size_t)XN * YN);
    return finish_plan(P, gpunum);
}

int plan_proj(ctlab_plan* P, const float* d_vol, float* d_prj, hipStream_t s)
{
    if (!P || !d_vol || !d_prj) {
        set_error(CTLAB_ERR_ARG, "NULL plan or buffer");
        return CTLAB_ERR_ARG;
    }
    if (!P->d_cz || !P->d_P) {
        set_error(CTLAB_ERR_ARG, "plan has no workspace (an earlier allocation failed)");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(P->device);
    if (is_blocked(P->in)) return blocked_proj(P, d_vol, d_prj, s);
    const DevGeom g = P->dev();
    hipError_t e = P->sat_bits == 32 ? launch_proj<float>(g, d_vol, P->d_cz, d_prj, s)
                                     : launch_proj<double>(g, d_vol, P->d_cz, d_prj, s);
    if (hip_check(e, "projector launch")) return t_err;
    return CTLAB_OK;
}

int plan_back(ctlab_plan* P, const float* d_prj, float* d_vol, hipStream_t s)
{
    if (!P || !d_vol || !d_prj) {
        set_error(CTLAB_ERR_ARG, "NULL plan or buffer");
        return CTLAB_ERR_ARG;
    }
    if (!P->d_cz || !P->d_P) {
        set_error(CTLAB_ERR_ARG, "plan has no workspace (an earlier allocation failed)");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(P->device);
    if (is_blocked(P->in)) return blocked_back(P, d_prj, d_vol, s);
    const DevGeom g = P->dev();
    const int ny = (int)P->views_y.size(), nx = (int)P->views_x.size();
    const int* vy = P->d_views;
    const int* vx = P->d_views + ny;
    // the integral-image workspace (>= XN*YN*ZN floats) is idle during backprojection
    float* scratch = static_cast<float*>(P->d_cz);
    if (P->d_start1) {
        hipError_t e = P->sat_bits == 32
                           ? launch_fan2d_back<float>(g, d_prj, P->d_P, d_vol, scratch, vy, ny, vx, nx,
                                                      P->d_start1, P->SJ, P->SR1, s)
                           : launch_fan2d_back<double>(g, d_prj, P->d_P, d_vol, scratch, vy, ny, vx, nx,
                                                       P->d_start1, P->SJ, P->SR1, s);
        if (hip_check(e, "backprojector launch")) return t_err;
        return CTLAB_OK;
    }
    hipError_t e = P->sat_bits == 32
                       ? launch_back<float>(g, d_prj, P->d_P, d_vol, scratch, vy, ny, vx, nx, P->d_rec, P->d_start,
                                            P->SJ, P->SR, s)
                       : launch_back<double>(g, d_prj, P->d_P, d_vol, scratch, vy, ny, vx, nx, P->d_rec,
                                             P->d_start, P->SJ, P->SR, s);
    if (hip_check(e, "backprojector launch")) return t_err;
    return CTLAB_OK;
}

// ---------------------------------------------------------------------------
// Plan cache of the host-pointer entry points.  The reference's callers call
// DD3Proj_gpu / DD3Back_gpu once per subset and iteration with the same
// geometry (src/DD3_GPU_recon.h:60-65); building the per-view tables, the
// integral-image workspace and the start table each time would cost more than
// the projection.  Each entry keeps a plan, device buffers for one volume and
// one projection set, and a stream, keyed by every geometry argument (compared
// byte for byte) and the device; least recently used entries are dropped
// beyond the limit.  An entry is used by one call at a time (its own lock);
// entries are never freed at process exit (the HIP runtime may be gone by
// then), only by ctlab_plan_cache_clear / a lower limit.
// ---------------------------------------------------------------------------
struct CacheEntry {
    std::vector<unsigned char> key;
    int device = 0;
    ctlab_plan* plan = nullptr;
    float* dv = nullptr;
    float* dp = nullptr;
    hipStream_t s = nullptr;
    // forward calls: the view set in consecutive chunks, each a plan on the
    // entry plan's workspace, so that a chunk's projections go back to the
    // host (on s2) while the next chunks are projected (on s)
    std::vector<ctlab_plan*> chunk;
    std::vector<int> chunk_p0;
    std::vector<hipEvent_t> chunk_done;
    hipStream_t s2 = nullptr;
    bool chunk_tried = false;
    std::mutex busy;
    ~CacheEntry()
    {
        if (!plan) return;
        DeviceGuard dg(device);
        if (s) hipStreamSynchronize(s);
        if (s2) hipStreamSynchronize(s2);
        for (ctlab_plan* c : chunk) free_plan(c);
        for (hipEvent_t e : chunk_done) hipEventDestroy(e);
        if (s2) hipStreamDestroy(s2);
        if (dv) hipFree(dv);
        if (dp) hipFree(dp);
        if (s) hipStreamDestroy(s);
        free_plan(plan);
    }
};

struct PlanCache {
    std::mutex mu;
    std::list<std::shared_ptr<CacheEntry>> lru;   // front: most recent
    int limit = 16;   // an OS-SART caller of the host entry points cycles through its subsets
};

static PlanCache& plan_cache()
{
    static PlanCache* c = new PlanCache();   // deliberately never destroyed
    return *c;
}

struct KeyWriter {
    std::vector<unsigned char> b;
    template <typename T> void put(const T& v)
    {
        const unsigned char* q = reinterpret_cast<const unsigned char*>(&v);
        b.insert(b.end(), q, q + sizeof(T));
    }
    template <typename T> void arr(const T* a, size_t n)
    {
        put(n);
        if (a && n) {
            const unsigned char* q = reinterpret_cast<const unsigned char*>(a);
            b.insert(b.end(), q, q + n * sizeof(T));
        } else {
            put((int)(a != nullptr));
        }
    }
};

// A plan and buffers for the key, from the cache or built by make(); null on
// error.  The entry comes back locked by `lock`.
template <class Make>
static std::shared_ptr<CacheEntry> cache_get(std::vector<unsigned char>&& key, int device, Make make,
                                             std::unique_lock<std::mutex>& lock)
{
    PlanCache& C = plan_cache();
    std::shared_ptr<CacheEntry> e;
    {
        std::lock_guard<std::mutex> g(C.mu);
        for (auto it = C.lru.begin(); it != C.lru.end(); ++it)
            if ((*it)->device == device && (*it)->key == key) {
                e = *it;
                C.lru.splice(C.lru.begin(), C.lru, it);
                break;
            }
    }
    if (e) {
        lock = std::unique_lock<std::mutex>(e->busy);
        return e;
    }
    // a new entry; a partly built one frees what it holds when dropped
    auto build = [&]() -> std::shared_ptr<CacheEntry> {
        auto n = std::make_shared<CacheEntry>();
        n->device = device;
        n->plan = make();
        if (!n->plan) return nullptr;
        DeviceGuard dg(device);
        const size_t nvox = (size_t)n->plan->in.XN * n->plan->in.YN * n->plan->in.ZN;
        const size_t nprj = (size_t)n->plan->in.PN * n->plan->in.DNU * n->plan->in.DNV;
        if (hip_check(hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking), "hipStreamCreate") ||
            hip_check(hipMalloc(&n->dv, nvox * sizeof(float)), "hipMalloc(volume)") ||
            hip_check(hipMalloc(&n->dp, nprj * sizeof(float)), "hipMalloc(projection)"))
            return nullptr;
        return n;
    };
    e = build();
    if (!e && t_err == CTLAB_ERR_NOMEM) {
        // Device memory ran out: the cached entries may hold it.  Take them
        // all out of the cache (an entry another thread is using is freed when
        // that call ends) and, if there were any, try once more.
        std::list<std::shared_ptr<CacheEntry>> dropped;
        {
            std::lock_guard<std::mutex> g(C.mu);
            dropped.swap(C.lru);
        }
        if (!dropped.empty()) {
            dropped.clear();
            clear_error();
            e = build();
        }
    }
    if (!e) return nullptr;
    e->key = std::move(key);
    lock = std::unique_lock<std::mutex>(e->busy);
    std::vector<std::shared_ptr<CacheEntry>> dropped;   // freed after the cache lock is released
    {
        std::lock_guard<std::mutex> g(C.mu);
        C.lru.push_front(e);
        while ((int)C.lru.size() > std::max(C.limit, 0)) {
            dropped.push_back(C.lru.back());
            C.lru.pop_back();
        }
    }
    return e;
}

// plan over the consecutive views [p0, p1) of P, borrowing P's workspace
static ctlab_plan* view_range_plan(const ctlab_plan* P, int p0, int p1)
{
    ctlab_plan* S = new ctlab_plan();
    S->in = P->in;
    S->in.PN = p1 - p0;
    S->in.angs.assign(P->in.angs.begin() + p0, P->in.angs.begin() + p1);
    S->in.zpos.assign(P->in.zpos.begin() + p0, P->in.zpos.begin() + p1);
    S->sat_bits = P->sat_bits;
    S->tile_nu = P->tile_nu;
    S->tile_nm = P->tile_nm;
    return finish_plan(S, P->device, P);
}

// Forward chunks of a cached entry: 4 chunks from 256 views, 2 from 128, none
// below or for blocked plans.  Each view's projection depends on that view
// only, so the chunked result equals the one-launch result bit for bit.  A
// failure leaves the entry on one launch and clears the error it set.
static void build_chunks(CacheEntry& E)
{
    E.chunk_tried = true;
    const ctlab_plan* P = E.plan;
    const int PN = P->in.PN;
    const int K = is_blocked(P->in) ? 1 : PN >= 256 ? 4 : PN >= 128 ? 2 : 1;
    if (K < 2) return;
    bool ok = !hip_check(hipStreamCreateWithFlags(&E.s2, hipStreamNonBlocking), "hipStreamCreate");
    for (int k = 0; ok && k < K; ++k) {
        const int p0 = (int)((long)PN * k / K), p1 = (int)((long)PN * (k + 1) / K);
        ctlab_plan* c = view_range_plan(P, p0, p1);
        hipEvent_t e = nullptr;
        ok = c != nullptr && !hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        if (c) E.chunk.push_back(c);
        if (e) E.chunk_done.push_back(e);
        E.chunk_p0.push_back(p0);
    }
    if (ok) {
        E.chunk_p0.push_back(PN);
        return;
    }
    for (ctlab_plan* c : E.chunk) free_plan(c);
    for (hipEvent_t e : E.chunk_done) hipEventDestroy(e);
    if (E.s2) hipStreamDestroy(E.s2);
    E.chunk.clear();
    E.chunk_done.clear();
    E.chunk_p0.clear();
    E.s2 = nullptr;
    clear_error();
}

// host-pointer round trip used by the CTLAB entry points, on a cached entry:
// H2D of the input, the operator, D2H of the output.  Forward calls with
// chunks project chunk k + 2 while chunk k's projections are copied back (a
// copy from device memory into pageable host memory can hold the calling
// thread until it completes, so two chunks are queued ahead of each copy).
static void host_call(CacheEntry& E, float* hvol, float* hprj, bool forward)
{
    ctlab_plan* P = E.plan;
    DeviceGuard dg(P->device);
    const size_t nvox = (size_t)P->in.XN * P->in.YN * P->in.ZN;
    const size_t nprj = (size_t)P->in.PN * P->in.DNU * P->in.DNV;
    bool ok;
    if (forward && !E.chunk_tried) build_chunks(E);
    if (forward && !E.chunk.empty()) {
        const int K = (int)E.chunk.size();
        const size_t pv = (size_t)P->in.DNU * P->in.DNV;
        auto d2h = [&](int k) {
            const size_t o = (size_t)E.chunk_p0[k] * pv, n = (size_t)(E.chunk_p0[k + 1] - E.chunk_p0[k]) * pv;
            return !hip_check(hipStreamWaitEvent(E.s2, E.chunk_done[k], 0), "stream wait") &&
                   !hip_check(hipMemcpyAsync(hprj + o, E.dp + o, n * sizeof(float), hipMemcpyDeviceToHost, E.s2),
                              "D2H");
        };
        ok = !hip_check(hipMemcpyAsync(E.dv, hvol, nvox * sizeof(float), hipMemcpyHostToDevice, E.s), "H2D");
        for (int k = 0; ok && k < K; ++k) {
            ok = plan_proj(E.chunk[k], E.dv, E.dp + (size_t)E.chunk_p0[k] * pv, E.s) == CTLAB_OK &&
                 !hip_check(hipEventRecord(E.chunk_done[k], E.s), "event record");
            if (ok && k >= 2) ok = d2h(k - 2);
        }
        for (int k = std::max(K - 2, 0); ok && k < K; ++k) ok = d2h(k);
        hipStreamSynchronize(E.s);
        if (ok) hip_check(hipStreamSynchronize(E.s2), "kernel execution");
        else hipStreamSynchronize(E.s2);
        return;
    }
    if (forward) {
        ok = !hip_check(hipMemcpyAsync(E.dv, hvol, nvox * sizeof(float), hipMemcpyHostToDevice, E.s), "H2D") &&
             plan_proj(P, E.dv, E.dp, E.s) == CTLAB_OK &&
             !hip_check(hipMemcpyAsync(hprj, E.dp, nprj * sizeof(float), hipMemcpyDeviceToHost, E.s), "D2H");
    } else {
        ok = !hip_check(hipMemcpyAsync(E.dp, hprj, nprj * sizeof(float), hipMemcpyHostToDevice, E.s), "H2D") &&
             plan_back(P, E.dp, E.dv, E.s) == CTLAB_OK &&
             !hip_check(hipMemcpyAsync(hvol, E.dv, nvox * sizeof(float), hipMemcpyDeviceToHost, E.s), "D2H");
    }
    if (ok) hip_check(hipStreamSynchronize(E.s), "kernel execution");
    else hipStreamSynchronize(E.s);
}

static std::vector<unsigned char> cone_key(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                                           const float* yds, const float* zds, float cx, float cy, float cz,
                                           const float* hangs, const float* hzPos, int PN, int XN, int YN, int ZN,
                                           float dx, float dz, const byte* mask)
{
    KeyWriter k;
    k.put(1);
    k.put(x0); k.put(y0); k.put(z0); k.put(DNU); k.put(DNV); k.put(cx); k.put(cy); k.put(cz);
    k.put(PN); k.put(XN); k.put(YN); k.put(ZN); k.put(dx); k.put(dz);
    const bool sizes_ok = DNU > 0 && DNV > 0 && PN > 0 && XN > 0 && YN > 0;
    k.arr(xds, sizes_ok ? (size_t)DNU : 0); k.arr(yds, sizes_ok ? (size_t)DNU : 0);
    k.arr(zds, sizes_ok ? (size_t)DNV : 0); k.arr(hangs, sizes_ok ? (size_t)PN : 0);
    k.arr(hzPos, sizes_ok ? (size_t)PN : 0); k.arr(mask, sizes_ok ? (size_t)XN * YN : 0);
    return std::move(k.b);
}

static void cone_host_call(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                           float cx, float cy, float cz, float* hangs, float* hzPos, int PN, int XN, int YN, int ZN,
                           float* hvol, float* hprj, float dx, float dz, byte* mask, int gpunum, bool forward)
{
    std::unique_lock<std::mutex> lock;
    auto e = cache_get(cone_key(x0, y0, z0, DNU, DNV, xds, yds, zds, cx, cy, cz, hangs, hzPos, PN, XN, YN, ZN, dx,
                                dz, mask),
                       gpunum,
                       [&] {
                           return plan_create_cone(x0, y0, z0, DNU, DNV, xds, yds, zds, cx, cy, cz, hangs, hzPos,
                                                   PN, XN, YN, ZN, dx, dz, mask, gpunum);
                       },
                       lock);
    if (e) host_call(*e, hvol, hprj, forward);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" {

int ctlab_last_error(void) { return ctlab::t_err; }
const char* ctlab_last_error_message(void) { return ctlab::t_msg; }

const char* ctlab_error_string(int code)
{
    switch (code) {
    case CTLAB_OK: return "success";
    case CTLAB_ERR_ARG: return "invalid argument";
    case CTLAB_ERR_MODE: return "unsupported projection mode";
    case CTLAB_ERR_GEOMETRY: return "unsupported geometry";
    case CTLAB_ERR_HIP: return "HIP runtime error";
    case CTLAB_ERR_NOMEM: return "device allocation failed";
    case CTLAB_ERR_IO: return "file input/output failed";
    default: return "unknown error";
    }
}

ctlab_plan* ctlab_plan_create(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                              const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                              float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                              int YN, int ZN, float dx, float dz, const byte* mask, int gpunum)
{
    clear_error();
    return plan_create_cone(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter,
                            hangs, hzPos, PN, XN, YN, ZN, dx, dz, mask, gpunum);
}

ctlab_plan* ctlab_plan_create_fan(float x0, float y0, const float* xds, const float* yds, int DNU,
                                  int SLN, float imgXCenter, float imgYCenter, int XN, int YN, float dx,
                                  const float* hangs, int PN, const byte* mask, int gpunum)
{
    clear_error();
    return plan_create_fan(x0, y0, xds, yds, DNU, SLN, imgXCenter, imgYCenter, XN, YN, dx, hangs, PN,
                           mask, gpunum);
}

void ctlab_plan_destroy(ctlab_plan* plan) { free_plan(plan); }

int ctlab_plan_set_sat_bits(ctlab_plan* plan, int bits)
{
    clear_error();
    if (!plan || (bits != 32 && bits != 64)) {
        set_error(CTLAB_ERR_ARG, "sat bits must be 32 or 64");
        return CTLAB_ERR_ARG;
    }
    if (plan->sat_bits == bits) return CTLAB_OK;
    DeviceGuard dg(plan->device);
    if (alloc_workspace(plan, bits)) return t_err;
    for (auto& pt : plan->parts) {   // the parts run on the blocked plan's workspace
        pt.plan->sat_bits = bits;
        borrow_workspace(pt.plan, plan);
    }
    return CTLAB_OK;
}

int ctlab_proj(ctlab_plan* plan, const float* d_vol, float* d_prj, void* stream)
{
    clear_error();
    return plan_proj(plan, d_vol, d_prj, (hipStream_t)stream);
}

int ctlab_back(ctlab_plan* plan, const float* d_prj, float* d_vol, void* stream)
{
    clear_error();
    return plan_back(plan, d_prj, d_vol, (hipStream_t)stream);
}

int ctlab_plan_set_tiles(ctlab_plan* plan, int nu, int nm)
{
    clear_error();
    auto ok = [](int t) { return t == 8 || t == 12 || t == 16; };
    if (!plan || !ok(nu) || !ok(nm)) {
        set_error(CTLAB_ERR_ARG, "tile sizes must be 8, 12 or 16");
        return CTLAB_ERR_ARG;
    }
    plan->tile_nu = nu;
    plan->tile_nm = nm;
    return CTLAB_OK;
}

int ctlab_plan_get_tiles(const ctlab_plan* plan, int* nu, int* nm)
{
    if (!plan) return CTLAB_ERR_ARG;
    if (nu) *nu = plan->tile_nu;
    if (nm) *nm = plan->tile_nm;
    return CTLAB_OK;
}

// Times every tile variant of both operators on the caller's buffers (each
// twice, the faster run counts) and keeps the fastest; d_prj and d_vol are
// overwritten.  ms_out (optional, 6 floats): proj 8/12/16, back 8/12/16.
int ctlab_plan_autotune(ctlab_plan* plan, float* d_vol, float* d_prj, float* ms_out)
{
    clear_error();
    if (!plan || !d_vol || !d_prj) {
        set_error(CTLAB_ERR_ARG, "NULL plan or buffer");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(plan->device);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") ||
        hip_check(hipEventCreate(&e0), "event") || hip_check(hipEventCreate(&e1), "event")) {
        if (s) hipStreamDestroy(s);
        if (e0) hipEventDestroy(e0);
        return ctlab_last_error();
    }
    const int tiles[3] = {8, 12, 16};
    float best[2] = {1e30f, 1e30f};
    int pick[2] = {plan->tile_nu, plan->tile_nm};
    int rc = CTLAB_OK;
    for (int op = 0; op < 2 && rc == CTLAB_OK; ++op) {
        for (int k = 0; k < 3 && rc == CTLAB_OK; ++k) {
            if (op == 0) plan->tile_nu = tiles[k]; else plan->tile_nm = tiles[k];
            float tmin = 1e30f;
            for (int rep = 0; rep < 2 && rc == CTLAB_OK; ++rep) {
                hipEventRecord(e0, s);
                rc = op == 0 ? plan_proj(plan, d_vol, d_prj, s) : plan_back(plan, d_prj, d_vol, s);
                hipEventRecord(e1, s);
                if (rc == CTLAB_OK && hip_check(hipEventSynchronize(e1), "autotune")) rc = ctlab_last_error();
                float ms = 0.0f;
                hipEventElapsedTime(&ms, e0, e1);
                if (ms < tmin) tmin = ms;
            }
            if (ms_out) ms_out[op * 3 + k] = tmin;
            if (tmin < best[op]) { best[op] = tmin; pick[op] = tiles[k]; }
        }
    }
    plan->tile_nu = pick[0];
    plan->tile_nm = pick[1];
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(s);
    return rc;
}

int ctlab_plan_views(const ctlab_plan* plan, int* n_yslab, int* n_xslab)
{
    if (!plan) return CTLAB_ERR_ARG;
    if (n_yslab) *n_yslab = (int)plan->views_y.size();
    if (n_xslab) *n_xslab = (int)plan->views_x.size();
    return CTLAB_OK;
}

int ctlab_plan_cache_set_limit(int entries)
{
    clear_error();
    if (entries < 0) {
        set_error(CTLAB_ERR_ARG, "cache limit must be >= 0");
        return -1;
    }
    PlanCache& C = plan_cache();
    std::vector<std::shared_ptr<CacheEntry>> dropped;   // freed after the cache lock is released
    int prev;
    {
        std::lock_guard<std::mutex> g(C.mu);
        prev = C.limit;
        C.limit = entries;
        while ((int)C.lru.size() > entries) {
            dropped.push_back(C.lru.back());
            C.lru.pop_back();
        }
    }
    return prev;
}

void ctlab_plan_cache_clear(void)
{
    const int prev = ctlab_plan_cache_set_limit(0);
    ctlab_plan_cache_set_limit(prev);
}

int ctlab_plan_cache_size(void)
{
    PlanCache& C = plan_cache();
    std::lock_guard<std::mutex> g(C.mu);
    return (int)C.lru.size();
}

long long ctlab_plan_workspace_bytes(const ctlab_plan* plan)
{
    if (!plan) return -1;
    return (long long)((plan->ws_cz_cap + plan->ws_P_cap) * plan->ws_elem);
}

void DD3Proj_gpu(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                 float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                 int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                 int gpunum, int prjMode)
{
    clear_error();
    VoidReport report("DD3Proj_gpu");
    if (prjMode != 0) {
        set_error(CTLAB_ERR_MODE, "prjMode %d not supported (branchless DD is mode 0)", prjMode);
        return;
    }
    if (!hvol || !hprj) {
        set_error(CTLAB_ERR_ARG, "NULL volume or projection pointer");
        return;
    }
    cone_host_call(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN, XN,
                   YN, ZN, hvol, hprj, dx, dz, mask, gpunum, true);
}

void DD3Back_gpu(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                 float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                 int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                 int gpunum, int squared, int prjMode)
{
    (void)squared;
    clear_error();
    VoidReport report("DD3Back_gpu");
    if (prjMode != 0) {
        set_error(CTLAB_ERR_MODE, "prjMode %d not supported (branchless DD is mode 0)", prjMode);
        return;
    }
    if (!hvol || !hprj) {
        set_error(CTLAB_ERR_ARG, "NULL volume or projection pointer");
        return;
    }
    cone_host_call(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN, XN,
                   YN, ZN, hvol, hprj, dx, dz, mask, gpunum, false);
}

void DD2_multiGPU(float* hvol, float* hprj, const int method, const float x0, const float y0, float* xds,
                  float* yds, const int DNU, const int SLN, const float imgXCenter, const float imgYCenter,
                  const int XN, const int YN, const float dx, float* hangs, int PN, byte* mask,
                  int* startIdx, const int gpuNum)
{
    clear_error();
    VoidReport report("DD2_multiGPU");
    if (method != 0 && method != 2) {
        set_error(CTLAB_ERR_MODE, "method %d not supported (0: branchless proj, 2: branchless back)", method);
        return;
    }
    if (!hvol || !hprj) {
        set_error(CTLAB_ERR_ARG, "NULL volume or projection pointer");
        return;
    }
    if (gpuNum < 1) {
        set_error(CTLAB_ERR_ARG, "gpuNum must be >= 1");
        return;
    }
    if (startIdx) {
        if (startIdx[0] != 0) {
            set_error(CTLAB_ERR_ARG, "startIdx[0] must be 0");
            return;
        }
        for (int i = 1; i < gpuNum; ++i)
            if (startIdx[i] < startIdx[i - 1] || startIdx[i] > SLN) {
                set_error(CTLAB_ERR_ARG, "startIdx must be non-decreasing within [0, SLN]");
                return;
            }
    }
    int dev = 0;
    hipGetDevice(&dev);
    KeyWriter k;
    k.put(0);
    k.put(x0); k.put(y0); k.put(DNU); k.put(SLN); k.put(imgXCenter); k.put(imgYCenter); k.put(XN); k.put(YN);
    k.put(dx); k.put(PN);
    const bool sizes_ok = DNU > 0 && SLN > 0 && PN > 0 && XN > 0 && YN > 0;
    k.arr(xds, sizes_ok ? (size_t)DNU : 0); k.arr(yds, sizes_ok ? (size_t)DNU : 0);
    k.arr(hangs, sizes_ok ? (size_t)PN : 0); k.arr(mask, sizes_ok ? (size_t)XN * YN : 0);
    std::unique_lock<std::mutex> lock;
    auto e = cache_get(std::move(k.b), dev,
                       [&] {
                           return plan_create_fan(x0, y0, xds, yds, DNU, SLN, imgXCenter, imgYCenter, XN, YN, dx,
                                                  hangs, PN, mask, dev);
                       },
                       lock);
    if (e) host_call(*e, hvol, hprj, method == 0);
}

// ---- multi-GPU split entry points ----------------------------------------
// The reference splits views (projector) or volume slices (backprojector)
// over GPUs 0..n-1 at the start indices given.  Every part is computed by the
// same operator, so the union equals one single-GPU call: here the whole view
// set runs on the current device, after the split is checked as the
// reference's callers build it (first index 0, non-decreasing, within range).

static bool split_ok(const int* start, int n, int limit, const char* what)
{
    if (n < 1 || !start) {
        set_error(CTLAB_ERR_ARG, "%s: need at least one GPU and its start index", what);
        return false;
    }
    if (start[0] != 0) {
        set_error(CTLAB_ERR_ARG, "%s[0] must be 0", what);
        return false;
    }
    for (int i = 1; i < n; ++i)
        if (start[i] < start[i - 1] || start[i] > limit) {
            set_error(CTLAB_ERR_ARG, "%s must be non-decreasing within [0, %d]", what, limit);
            return false;
        }
    return true;
}

static int current_device()
{
    int dev = 0;
    hipGetDevice(&dev);
    return dev;
}

void DD3Proj_multiGPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                      float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                      int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                      int prjMode, int* startPN, int gpuNum)
{
    clear_error();
    VoidReport report("DD3Proj_multiGPU");
    if (prjMode != 0) {
        set_error(CTLAB_ERR_MODE, "prjMode %d not supported (branchless DD is mode 0)", prjMode);
        return;
    }
    if (!split_ok(startPN, gpuNum, PN, "startPN")) return;
    DD3Proj_gpu(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN, XN,
                YN, ZN, hvol, hprj, dx, dz, mask, current_device(), prjMode);
}

void DD3Back_multiGPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                      float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                      int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                      int methodId, int* startVOL, int gpuNum)
{
    clear_error();
    VoidReport report("DD3Back_multiGPU");
    if (methodId != 0) {
        set_error(CTLAB_ERR_MODE, "methodId %d not supported (branchless DD is mode 0)", methodId);
        return;
    }
    if (!split_ok(startVOL, gpuNum, ZN, "startVOL")) return;
    DD3Back_gpu(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN, XN,
                YN, ZN, hvol, hprj, dx, dz, mask, current_device(), 0, methodId);
}

void DD3ProjHelical_3GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                         float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                         int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                         int methodId, int* startPN)
{
    DD3Proj_multiGPU(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN,
                     XN, YN, ZN, hvol, hprj, dx, dz, mask, methodId, startPN, 3);
}

void DD3ProjHelical_4GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                         float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                         int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                         int methodId, int* startPN)
{
    DD3Proj_multiGPU(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN,
                     XN, YN, ZN, hvol, hprj, dx, dz, mask, methodId, startPN, 4);
}

void DD3BackHelical_3GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                         float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                         int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                         int methodId, int* startVOL)
{
    DD3Back_multiGPU(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN,
                     XN, YN, ZN, hvol, hprj, dx, dz, mask, methodId, startVOL, 3);
}

void DD3BackHelical_4GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                         float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos, int PN,
                         int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz, byte* mask,
                         int methodId, int* startVOL)
{
    DD3Back_multiGPU(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter, imgZCenter, hangs, hzPos, PN,
                     XN, YN, ZN, hvol, hprj, dx, dz, mask, methodId, startVOL, 4);
}

}  // extern "C"

// Internal plan object shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dd_common.hpp"

namespace ctlab {

struct GeomIn {
    double x0, y0, z0;
    int DNU, DNV;
    std::vector<double> xds, yds, zds, angs, zpos;
    double cx, cy, cz;
    int PN, XN, YN, ZN;
    double dx, dz;
    std::vector<unsigned char> mask;  // empty = none
    int cone;
    std::vector<double> zbnd;         // DNV+1 row boundaries when given (row blocks of a larger
                                      // detector); empty: DD3Boundaries of zds
};

// The kernels hold one voxel column (z) and one detector column (v) in the 64
// lanes of a wavefront.  Larger volumes and detectors are split into blocks of
// at most kBlock slices and kBlock rows; a plan over them is the sum of one
// sub-plan per (slice block, row block) pair that rays can join.
constexpr int kBlock = 64;
inline bool is_blocked(const GeomIn& in) { return in.ZN > kBlock || in.DNV > kBlock; }

// fan beam with at most kFanSlices slices (ZN = DNV): lanes over detector
// boundaries (projector) and over row boundaries (backprojector)
constexpr int kFanSlices = 16;
constexpr int kFanProjSlices = 8;   // the projector's bound (see fan2d.hip)
bool is_fan2d(const DevGeom& g);
inline bool is_fan2d(const GeomIn& in) { return !in.cone && in.ZN == in.DNV && in.ZN <= kFanSlices; }

// integral-image elements: one per voxel (the projector's row reads stay on
// the rows of the slab it walks); a fan beam of few slices holds instead the
// row prefixes of both slab orientations per slice (YN * (XN + 1) + XN *
// (YN + 1) entries each)
inline size_t cz_elems(const GeomIn& in)
{
    const size_t vox = (size_t)in.XN * in.YN * in.ZN;
    if (!is_fan2d(in)) return vox;
    return std::max(vox, in.ZN * ((size_t)in.YN * (in.XN + 1) + (size_t)in.XN * (in.YN + 1)));
}
template <typename T>
hipError_t launch_fan2d_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s);
// d_Q: ZN * PN * (DNU + 1) elements of work space (the columns' prefix per slice and view)
template <typename T>
hipError_t launch_fan2d_back(const DevGeom& g, const float* d_prj, void* d_Q, float* d_vol, float* d_scratch,
                             const int* d_views_y, int ny, const int* d_views_x, int nx,
                             const unsigned short* d_start1, int SJ, int SR1, hipStream_t s);
// walk-order column containing each row boundary r = 0 .. SR1 - 1 of each (view, slab)
hipError_t launch_fan2d_start(const DevGeom& g, const void* d_rec, unsigned short* d_start1, int SJ, int SR1,
                              hipStream_t s);


template <typename T>
hipError_t launch_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s);
// dst[r*nd + k] = src[r*ns + off + k] for rows r < nrows, k < nd
hipError_t launch_rows_gather(const float* src, float* dst, long nrows, int ns, int off, int nd, hipStream_t s);
// dst[r*nd + off + k] = (add ? dst[...] : 0) + (src ? src[r*ns + k] : 0) for r < nrows, k < ns
hipError_t launch_rows_scatter(const float* src, float* dst, long nrows, int ns, int nd, int off, int add,
                               hipStream_t s);
// d_scratch: XN*YN*ZN floats of work space (the x-slab part before it is added)
template <typename T>
hipError_t launch_back(const DevGeom& g, const float* d_prj, void* d_P, float* d_vol, float* d_scratch,
                       const int* d_views_y, int ny, const int* d_views_x, int nx, const void* d_rec,
                       const unsigned short* d_start, int SJ, int SR, hipStream_t s);
hipError_t launch_back_start(const DevGeom& g, const void* d_rec, unsigned short* d_start, int SJ, int SR,
                             hipStream_t s);

}  // namespace ctlab

struct ctlab_plan {
    int device = 0;
    ctlab::GeomIn in;
    std::vector<ctlab::ViewParams> vp;
    std::vector<double> kap, tau, colfac, exy2, invdk, esdz, zhat, ez, rowfac, invs;
    double zhat0 = 0, dzh = 1;
    std::vector<int> views_y, views_x;
    // device copies
    ctlab::ViewParams* d_vp = nullptr;
    double *d_kap = nullptr, *d_tau = nullptr, *d_colfac = nullptr, *d_exy2 = nullptr;
    double *d_invdk = nullptr, *d_esdz = nullptr, *d_invs = nullptr;
    double *d_zhat = nullptr, *d_ez = nullptr, *d_rowfac = nullptr;
    unsigned char* d_mask = nullptr;
    int* d_views = nullptr;
    std::vector<double> rec;            // (PN*(DNU+1)+1)*4: walk-order column records (sign(s0)*klo, sign(s0)*khi, 1/|dk|, esdz), sentinel per view
    double* d_rec = nullptr;
    unsigned short* d_start = nullptr;  // PN*SJ*SR first walk-order column per (view, slab, 8 rows)
    int SJ = 0, SR = 0;
    unsigned short* d_start1 = nullptr; // fan beam of few slices: PN*SJ*SR1, the column per row boundary
    int SR1 = 0;
    void* d_cz = nullptr;   // z-prefix integral image (nvox elements)
    void* d_P = nullptr;    // projection v-prefix (nprj elements)
    int sat_bits = 64;
    size_t ws_elem = 0;
    size_t ws_cz_cap = 0, ws_P_cap = 0;   // workspace capacity in elements
    bool ws_shared = false; // d_cz / d_P belong to another plan (subset plans of a reconstruction)
    int tile_nu = 8, tile_nm = 16;     // the autotuner's choice at 512x512x64, 984 views

    // Blocked plans (ZN or DNV above kBlock): no tables of their own; parts[i]
    // covers slices [z0, z0+nz) and rows [v0, v0+nv).  All parts run on one
    // workspace (the first part's) and on the staging buffers below.
    struct Part {
        ctlab_plan* plan;
        int zb, vb;          // block indices
        int z0, nz, v0, nv;
    };
    std::vector<Part> parts;
    float* d_sub_vol = nullptr;   // XN*YN*min(ZN, kBlock): one slice block
    float* d_sub_prj = nullptr;   // PN*DNU*min(DNV, kBlock): one row block
    bool sub_shared = false;      // staging buffers belong to another plan

    ctlab::DevGeom dev() const
    {
        ctlab::DevGeom g;
        g.XN = in.XN; g.YN = in.YN; g.ZN = in.ZN; g.DNU = in.DNU; g.DNV = in.DNV; g.PN = in.PN;
        g.dx = in.dx; g.dz = in.dz; g.cone = in.cone;
        g.vp = ctlab::gmem(d_vp); g.kap = ctlab::gmem(d_kap); g.tau = ctlab::gmem(d_tau);
        g.colfac = ctlab::gmem(d_colfac); g.exy2 = ctlab::gmem(d_exy2);
        g.invdk = ctlab::gmem(d_invdk); g.esdz = ctlab::gmem(d_esdz);
        g.zhat = ctlab::gmem(d_zhat); g.ez = ctlab::gmem(d_ez); g.rowfac = ctlab::gmem(d_rowfac);
        g.mask = in.mask.empty() ? nullptr : ctlab::gmem(d_mask);
        g.invs = ctlab::gmem(d_invs);
        g.sjs = in.XN > in.YN ? in.XN : in.YN;
        g.zhat0 = zhat0; g.dzh = dzh;
        g.tile_nu = tile_nu; g.tile_nm = tile_nm;
        return g;
    }
};

namespace ctlab {
// builds the device tables of a plan whose `in` is filled; frees P and returns null on error.
// With share_ws, the plan borrows that plan's integral-image / prefix workspace
// (same SAT precision, at least as large); the two must then run in sequence.
ctlab_plan* finish_plan(ctlab_plan* P, int gpunum, const ctlab_plan* share_ws = nullptr);
void free_plan(ctlab_plan* P);
int plan_proj(ctlab_plan* P, const float* d_vol, float* d_prj, hipStream_t s);
int plan_back(ctlab_plan* P, const float* d_prj, float* d_vol, hipStream_t s);

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        hipGetDevice(&prev);
        if (prev != dev) hipSetDevice(dev); else prev = -1;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) hipSetDevice(prev);
    }
};
}  // namespace ctlab

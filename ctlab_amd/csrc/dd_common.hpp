// Shared host/device types of the MI355X branchless distance-driven library.
//
// Conventions (identical in the CPU oracle):
//   vol[(y*XN + x)*ZN + z]  (the reference's (Z,X,Y) order, src/utilities.cuh:441-454)
//   prj[(p*DNU + u)*DNV + v]  (v fastest, mirrors the z-fastest volume)
//   mask[y*XN + x]
// For each view the slabs are the voxel planes perpendicular to the in-plane
// axis the central ray runs along ("slab axis"); inside a slab, voxel rows run
// along the other in-plane axis ("row axis") and voxels along z.  A detector
// cell's footprint on slab j is the rectangle [a_u, a_u+1] x [b_uv, b_uv+1] in
// lattice units, a = row coordinate, b = z coordinate.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

// Device pointers in the structs below address global memory.  The device pass
// says so in their types: a kernel that reads such a pointer out of a
// descriptor array (not a kernel argument) would otherwise get flat loads,
// which also count on the LDS wait counter and would serialise the walk
// loops' LDS reads behind their prefetched memory loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define CTLAB_GMEM __attribute__((address_space(1)))
#else
#define CTLAB_GMEM
#endif

namespace ctlab {

// generic -> global pointer (explicit address-space cast in the device pass)
template <typename T>
__host__ __device__ inline CTLAB_GMEM T* gmem(T* p) { return (CTLAB_GMEM T*)p; }

// Per-view constants, computed on the host in FP64 (plan build).
struct ViewParams {
    double s0;        // centre of slab 0 minus source slab-axis coordinate (s_j - S = s0 + j*dx)
    double rbase;     // (source row-axis coordinate - row-axis lower edge) / dx
    double beta0;     // (source z - Zmin) / dz ; 0 for multi-slice fan beam
    int yslab;        // 1: slabs are y = const planes, rows run along x
    int NS, NR;       // slab count, rows per slab
    int slab_stride;  // z-column stride between slabs
    int row_stride;   // z-column stride between rows
    int asc;          // 1 if the row coordinate a grows with detector index u
    int top_total;    // 1 if, for every slab and column, the top voxel boundary z = ZN maps
                      // at or above the last detector row boundary (backprojector shortcut)
};

// Everything a kernel needs, passed by value.
struct DevGeom {
    int XN, YN, ZN, DNU, DNV, PN;
    double dx, dz;
    int cone;                 // 1: cone beam, 0: multi-slice fan beam
    const CTLAB_GMEM ViewParams* vp;     // PN
    const CTLAB_GMEM double* kap;        // PN*(DNU+1): d(row)/d(slab) slope of each u-boundary ray
    const CTLAB_GMEM double* tau;        // PN*DNU: 1 / (cell-centre slab coord - source slab coord)
    const CTLAB_GMEM double* colfac;     // PN*DNU: per-column output factor
    const CTLAB_GMEM double* exy2;       // PN*DNU: squared in-plane source->cell-centre distance
    const CTLAB_GMEM double* invdk;      // PN*DNU: 1/|kap[u+1]-kap[u]|
    const CTLAB_GMEM double* esdz;       // PN*DNU: (cell-centre slab coord - source slab coord) / dzh
    const CTLAB_GMEM double* zhat;       // DNV+1: row boundaries, (z - z0)/dz (cone) or v (fan)
    const CTLAB_GMEM double* ez;         // DNV: row centre z - z0 (0 for fan beam)
    const CTLAB_GMEM double* rowfac;     // DNV: 1/(zhat[v+1]-zhat[v]) (cone) or 1 (fan)
    const CTLAB_GMEM unsigned char* mask;// XN*YN or nullptr
    const CTLAB_GMEM double* invs;  // PN*sjs: 1 / (slab centre - source slab coordinate), slab j of view p
    int sjs;                  // row stride of invs (max(XN, YN))
    double zhat0, dzh;        // uniform row model used by the backprojector
    int tile_nu;              // projector: detector columns per wavefront (8, 12, 16)
    int tile_nm;              // backprojector: voxel rows per wavefront (8, 12, 16)
};

}  // namespace ctlab

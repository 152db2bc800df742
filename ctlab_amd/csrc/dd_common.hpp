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

namespace ctlab {

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
};

// Everything a kernel needs, passed by value.
struct DevGeom {
    int XN, YN, ZN, DNU, DNV, PN;
    double dx, dz;
    int cone;                 // 1: cone beam, 0: multi-slice fan beam
    const ViewParams* vp;     // PN
    const double* kap;        // PN*(DNU+1): d(row)/d(slab) slope of each u-boundary ray
    const double* tau;        // PN*DNU: 1 / (cell-centre slab coord - source slab coord)
    const double* colfac;     // PN*DNU: per-column output factor
    const double* exy2;       // PN*DNU: squared in-plane source->cell-centre distance
    const double* invdk;      // PN*DNU: 1/|kap[u+1]-kap[u]|
    const double* esdz;       // PN*DNU: (cell-centre slab coord - source slab coord) / dzh
    const double* zhat;       // DNV+1: row boundaries, (z - z0)/dz (cone) or v (fan)
    const double* ez;         // DNV: row centre z - z0 (0 for fan beam)
    const double* rowfac;     // DNV: 1/(zhat[v+1]-zhat[v]) (cone) or 1 (fan)
    const unsigned char* mask;// XN*YN or nullptr
    double zhat0, dzh;        // uniform row model used by the backprojector
    int tile_nu;              // projector: detector columns per wavefront (8, 16, 32)
    int tile_nm;              // backprojector: voxel rows per wavefront (8, 16, 32)
};

}  // namespace ctlab

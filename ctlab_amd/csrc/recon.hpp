// Internal reconstruction object shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "plan.hpp"

struct ctlab_recon {
    int device = 0;
    int osNum = 1;
    long nvox = 0, nprj = 0;
    int XN = 0, YN = 0, ZN = 0;
    std::vector<ctlab_plan*> sub;       // subset plans (interleaved views)
    std::vector<long> sub_off;          // element offset of each subset in the permuted projections
    std::vector<int> perm;              // permuted position -> original view
    ctlab_plan* full = nullptr;         // all views, original order (FISTA / CG)
    float* d_b = nullptr;               // measured projections, original order
    float* d_bperm = nullptr;           // measured projections, subset order
    float* d_rowsum = nullptr;          // per subset A_s * 1 (subset order)
    float* d_colsum = nullptr;          // per subset A_s^T * 1 (osNum volumes)
    float* d_prj = nullptr;             // projection work buffer (nprj)
    float* d_vol = nullptr;             // volume work buffers
    float* d_vol2 = nullptr;
    float* d_vol3 = nullptr;
    float* d_vol4 = nullptr;
    float* d_tv = nullptr;              // 9 * nvox for the FGP prox
    double* d_part = nullptr;           // reduction partials
    double* d_red = nullptr;            // reduction results (4 doubles)
    bool have_data = false;
    hipGraphExec_t sart_graph = nullptr;
    hipStream_t graph_stream = nullptr;
    float graph_lambda = 0.0f;
    int graph_nonneg = -1;
    float* graph_x = nullptr;           // volume the captured graph updates in place
};

#include <functional>

namespace ctlab {
// CG on the normal equations (the body of ctlab_cg); after_iter(k, d_x), when
// set, runs after iteration k (1-based) and stops the run by returning non-zero.
// When stop is set, ||Ax - b||^2 is carried along (each step lowers it by
// 2 alpha <r,d> - alpha^2 ||Ad||^2), written to stop->h_fid[k-1] (or not, if
// NULL), and the run ends after the first step that takes it to f_target or
// below; stop->done receives the number of steps taken.
struct CgStop {
    double f_target;
    double* h_fid;
    int done;
};
int cg_run(ctlab_recon* R, float* d_x, int iters, int cgm, double* h_resid, hipStream_t s,
           const std::function<int(int, const float*)>& after_iter, CgStop* stop = nullptr);
}  // namespace ctlab

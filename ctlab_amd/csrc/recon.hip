// Device-resident reconstruction drivers over a projector plan.
//
//   OS-SART ... src/DD3_GPU_recon.h:30-65 (OS_SART), weighting src/SARTWeighting.h:22-91,
//               row / column sums A*1 and A^T*1 as evalProjBack (src/DD3_GPU_recon.h:21-28)
//   CG ........ src/DD3_GPU_recon.h:67-103 (CG, CGMethod :14-19), updates
//               src/utilities.hpp:1121-1156, beta src/utilities.hpp:1228-1268, 1611-1621
//   FISTA-TV .. momentum src/FastIterativeShrinkageAlgorithm.h:17-42; the TV prox (FGP)
//               is this library's (the reference's TV routines have no bodies)
// The reference drivers copy every subset between host and device; here the
// data, the subset plans, the row/column sums and all work buffers stay on the
// GPU, and one OS-SART iteration (all subsets) is captured once in a HIP graph
// and replayed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/ctlab_mi355x.h"
#include "plan.hpp"
#include "recon.hpp"
#include "status.hpp"

namespace ctlab {
hipError_t rk_sart_prj_weight(float* prj, const float* real, const float* rowSum, long n, hipStream_t s);
hipError_t rk_sart_bak_weight(const float* vol, float* recon, const float* colSum, float lambda, int clamp_lo,
                              float lo, long n, hipStream_t s);
hipError_t rk_ossart_update(float* img, const float* cor, const float* weg, const float* msk, float lambda,
                            long n, hipStream_t s);
hipError_t rk_fista(float* cur, const float* las, float t1, float t2, float minV, float maxV, long n,
                    hipStream_t s);
hipError_t rk_soft_threshold(float* x, float g, long n, hipStream_t s);
hipError_t rk_axpby(float a, const float* x, float b, float* y, long n, hipStream_t s);
hipError_t rk_xpaymz(float* out, const float* x, float a, const float* y, const float* z, long n, hipStream_t s);
hipError_t rk_fill(float* x, float v, long n, hipStream_t s);
hipError_t rk_beta(const float* r0, const float* r1, const float* d, int cases, int from_residual, long n,
                   double* part, double* out, hipStream_t s);
hipError_t rk_dot(const float* x, const float* y, long n, double* part, double* out, hipStream_t s);
int red_scratch_doubles();
hipError_t rk_tv_prox(const float* b, float* x, float lam, int iters, int XN, int YN, int ZN, float* work,
                      hipStream_t s);
}  // namespace ctlab

using namespace ctlab;


namespace {

int chk(hipError_t e, const char* what) { return hip_check(e, what); }

void destroy(ctlab_recon* R)
{
    if (!R) return;
    DeviceGuard dg(R->device);
    for (auto* p : R->sub) free_plan(p);
    if (R->full) free_plan(R->full);
    float* fb[] = {R->d_b, R->d_bperm, R->d_rowsum, R->d_colsum, R->d_prj, R->d_vol, R->d_vol2, R->d_vol3,
                   R->d_vol4, R->d_tv};
    for (float* q : fb)
        if (q) hipFree(q);
    if (R->d_part) hipFree(R->d_part);
    if (R->d_red) hipFree(R->d_red);
    if (R->sart_graph) hipGraphExecDestroy(R->sart_graph);
    if (R->graph_stream) hipStreamDestroy(R->graph_stream);
    delete R;
}

ctlab_plan* subset_plan(const ctlab_plan* P, const std::vector<int>& views, const ctlab_plan* share_ws)
{
    ctlab_plan* S = new ctlab_plan();
    S->in = P->in;
    S->in.PN = (int)views.size();
    S->in.angs.resize(views.size());
    S->in.zpos.resize(views.size());
    for (size_t k = 0; k < views.size(); ++k) {
        S->in.angs[k] = P->in.angs[views[k]];
        S->in.zpos[k] = P->in.zpos[views[k]];
    }
    S->sat_bits = P->sat_bits;
    S->tile_nu = P->tile_nu;
    S->tile_nm = P->tile_nm;
    return finish_plan(S, P->device, share_ws);
}

double read_d(const double* d, hipStream_t s, int idx = 0)
{
    double h[4] = {0, 0, 0, 0};
    if (chk(hipMemcpyAsync(h, d, sizeof(double) * (idx + 1), hipMemcpyDeviceToHost, s), "D2H")) return NAN;
    if (chk(hipStreamSynchronize(s), "sync")) return NAN;
    return h[idx];
}

// <x, y> in FP64 (y null: ||x||^2), synchronous
double dot(ctlab_recon* R, const float* x, const float* y, long n, hipStream_t s)
{
    if (chk(rk_dot(x, y, n, R->d_part, R->d_red, s), "dot")) return NAN;
    return read_d(R->d_red, s);
}

// one OS-SART pass over all subsets (stream-ordered, capturable)
int sart_pass(ctlab_recon* R, float* d_x, float lambda, int nonneg, hipStream_t s)
{
    for (int k = 0; k < R->osNum; ++k) {
        ctlab_plan* P = R->sub[k];
        const long np = (long)P->in.PN * P->in.DNU * P->in.DNV;
        if (plan_proj(P, d_x, R->d_prj, s)) return -1;
        if (chk(rk_sart_prj_weight(R->d_prj, R->d_bperm + R->sub_off[k], R->d_rowsum + R->sub_off[k], np, s),
                "sart weighting")) return -1;
        if (plan_back(P, R->d_prj, R->d_vol, s)) return -1;
        if (chk(rk_sart_bak_weight(R->d_vol, d_x, R->d_colsum + (long)k * R->nvox, lambda, nonneg, 0.0f,
                                   R->nvox, s), "sart update")) return -1;
    }
    return 0;
}

// r = A x - b (original view order), returns ||r||^2 when want_norm
double residual(ctlab_recon* R, const float* d_x, hipStream_t s, bool want_norm)
{
    if (plan_proj(R->full, d_x, R->d_prj, s)) return NAN;
    if (chk(rk_xpaymz(R->d_prj, R->d_prj, -1.0f, R->d_b, nullptr, R->nprj, s), "residual")) return NAN;
    // d_prj = prj - b  (xpaymz: out = x + a*(y - z) with a = -1, y = b, z = 0)
    return want_norm ? dot(R, R->d_prj, nullptr, R->nprj, s) : 0.0;
}

}  // namespace

namespace ctlab {
// CG on the normal equations (body of ctlab_cg); after_iter(k, d_x), when
// set, runs after iteration k and stops the run by returning non-zero
int cg_run(ctlab_recon* R, float* d_x, int iters, int cgm, double* h_resid, hipStream_t s,
           const std::function<int(int, const float*)>& after_iter, CgStop* stop)
{
    if (!R || !d_x || iters < 0 || cgm < 0 || cgm > 4 || !R->have_data) {
        set_error(CTLAB_ERR_ARG, "need recon with data, volume, iters >= 0 and 0 <= cgm <= 4");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(R->device);
    void* stream = (void*)s;
    const long n = R->nvox;
    float* r = R->d_vol2;
    float* rn = R->d_vol3;
    float* d = R->d_vol4;
    float* w = R->d_vol;
    // r = A^T (b - A x) = -gradient; with a stop target that x already meets,
    // only the residual is formed and no step is taken
    double f;
    if (stop) {
        stop->done = 0;
        f = residual(R, d_x, s, true);
        if (std::isnan(f)) return ctlab_last_error();
        if (f <= stop->f_target) return chk(hipStreamSynchronize(s), "cg") ? ctlab_last_error() : CTLAB_OK;
        if (plan_back(R->full, R->d_prj, r, s)) return ctlab_last_error();
    } else {
        f = ctlab_recon_gradient(R, d_x, r, stream);
        if (std::isnan(f)) return ctlab_last_error();
    }
    const int steps = iters;
    if (chk(rk_axpby(0.0f, r, -1.0f, r, n, s), "negate")) return ctlab_last_error();
    hipMemcpyAsync(d, r, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    for (int it = 0; it < steps; ++it) {
        if (plan_proj(R->full, d, R->d_prj, s)) return ctlab_last_error();
        const double qq = dot(R, R->d_prj, nullptr, R->nprj, s);
        const double rd = dot(R, r, d, n, s);
        if (!(qq > 0)) break;
        const float alpha = (float)(rd / qq);
        if (chk(rk_axpby(alpha, d, 1.0f, d_x, n, s), "update1")) return ctlab_last_error();   // X + alpha D
        if (stop) {
            // ||A(x + alpha d) - b||^2 = f - 2 alpha <A^T(b - Ax), d> + alpha^2 ||Ad||^2
            f += (double)alpha * ((double)alpha * qq - 2.0 * rd);
            if (stop->h_fid) stop->h_fid[it] = f;
            stop->done = it + 1;
            if (f <= stop->f_target) break;
        }
        if (plan_back(R->full, R->d_prj, w, s)) return ctlab_last_error();                     // tem2 = A^T A d
        hipMemcpyAsync(rn, r, n * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (chk(rk_axpby(-alpha, w, 1.0f, rn, n, s), "update2")) return ctlab_last_error();   // R - alpha tem2
        if (chk(rk_beta(r, rn, d, cgm, 1, n, R->d_part, R->d_red, s), "beta")) return ctlab_last_error();
        double nd[2];
        if (chk(hipMemcpyAsync(nd, R->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H") ||
            chk(hipStreamSynchronize(s), "sync"))
            return ctlab_last_error();
        const float beta = nd[1] != 0.0 ? (float)(nd[0] / nd[1]) : 0.0f;
        if (chk(rk_axpby(1.0f, rn, beta, d, n, s), "update3")) return ctlab_last_error();     // R + beta D
        float* tmp = r; r = rn; rn = tmp;
        if (h_resid) h_resid[it] = dot(R, r, nullptr, n, s);
        if (after_iter && after_iter(it + 1, d_x)) return ctlab_last_error();
    }
    return chk(hipStreamSynchronize(s), "cg") ? ctlab_last_error() : CTLAB_OK;
}
}  // namespace ctlab

extern "C" {

ctlab_recon* ctlab_recon_create(const ctlab_plan* plan, int osNum)
{
    clear_error();
    if (!plan || osNum < 1 || osNum > plan->in.PN) {
        set_error(CTLAB_ERR_ARG, "need a plan and 1 <= osNum <= PN");
        return nullptr;
    }
    DeviceGuard dg(plan->device);
    ctlab_recon* R = new ctlab_recon();
    R->device = plan->device;
    R->osNum = osNum;
    R->XN = plan->in.XN; R->YN = plan->in.YN; R->ZN = plan->in.ZN;
    R->nvox = (long)R->XN * R->YN * R->ZN;
    const long cell = (long)plan->in.DNU * plan->in.DNV;
    R->nprj = (long)plan->in.PN * cell;
    std::vector<int> all(plan->in.PN);
    for (int p = 0; p < plan->in.PN; ++p) all[p] = p;
    R->full = subset_plan(plan, all, nullptr);
    if (!R->full) { destroy(R); return nullptr; }
    // the subset plans run one at a time: they borrow the full plan's workspace
    long off = 0;
    for (int k = 0; k < osNum; ++k) {
        std::vector<int> views;
        for (int p = k; p < plan->in.PN; p += osNum) views.push_back(p);
        ctlab_plan* S = subset_plan(plan, views, R->full);
        if (!S) { destroy(R); return nullptr; }
        R->sub.push_back(S);
        R->sub_off.push_back(off);
        off += (long)views.size() * cell;
        R->perm.insert(R->perm.end(), views.begin(), views.end());
    }
    const size_t fb = sizeof(float);
    if (chk(hipMalloc(&R->d_b, R->nprj * fb), "hipMalloc") || chk(hipMalloc(&R->d_bperm, R->nprj * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_rowsum, R->nprj * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_colsum, (size_t)osNum * R->nvox * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_prj, R->nprj * fb), "hipMalloc") || chk(hipMalloc(&R->d_vol, R->nvox * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_vol2, R->nvox * fb), "hipMalloc") || chk(hipMalloc(&R->d_vol3, R->nvox * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_vol4, R->nvox * fb), "hipMalloc") ||
        chk(hipMalloc(&R->d_part, sizeof(double) * red_scratch_doubles()), "hipMalloc") ||
        chk(hipMalloc(&R->d_red, sizeof(double) * 4), "hipMalloc")) {
        destroy(R);
        return nullptr;
    }
    // row sums A_s 1 and column sums A_s^T 1 (evalProjBack with all-ones inputs)
    hipStream_t s = nullptr;
    if (chk(rk_fill(R->d_vol, 1.0f, R->nvox, s), "fill")) { destroy(R); return nullptr; }
    for (int k = 0; k < osNum; ++k) {
        ctlab_plan* P = R->sub[k];
        const long np = (long)P->in.PN * cell;
        if (plan_proj(P, R->d_vol, R->d_rowsum + R->sub_off[k], s) ||
            chk(rk_fill(R->d_prj, 1.0f, np, s), "fill") ||
            plan_back(P, R->d_prj, R->d_colsum + (long)k * R->nvox, s)) {
            destroy(R);
            return nullptr;
        }
    }
    if (chk(hipDeviceSynchronize(), "row/column sums")) { destroy(R); return nullptr; }
    return R;
}

void ctlab_recon_destroy(ctlab_recon* R) { destroy(R); }

int ctlab_recon_set_tiles(ctlab_recon* R, int nu, int nm)
{
    clear_error();
    auto ok = [](int t) { return t == 8 || t == 12 || t == 16; };
    if (!R || !ok(nu) || !ok(nm)) {
        set_error(CTLAB_ERR_ARG, "tile sizes must be 8, 12 or 16");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(R->device);
    for (ctlab_plan* P : R->sub) { P->tile_nu = nu; P->tile_nm = nm; }
    R->full->tile_nu = nu;
    R->full->tile_nm = nm;
    if (R->sart_graph) {   // captured with the old kernels
        if (chk(hipStreamSynchronize(R->graph_stream), "sync")) return ctlab_last_error();
        hipGraphExecDestroy(R->sart_graph);
        R->sart_graph = nullptr;
    }
    return CTLAB_OK;
}

int ctlab_recon_set_data(ctlab_recon* R, const float* d_b, void* stream)
{
    clear_error();
    if (!R || !d_b) {
        set_error(CTLAB_ERR_ARG, "NULL recon or data");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(R->device);
    hipStream_t s = (hipStream_t)stream;
    const long cell = (long)R->full->in.DNU * R->full->in.DNV;
    if (chk(hipMemcpyAsync(R->d_b, d_b, R->nprj * sizeof(float), hipMemcpyDeviceToDevice, s), "D2D"))
        return ctlab_last_error();
    for (size_t q = 0; q < R->perm.size(); ++q)
        if (chk(hipMemcpyAsync(R->d_bperm + (long)q * cell, d_b + (long)R->perm[q] * cell, cell * sizeof(float),
                               hipMemcpyDeviceToDevice, s), "D2D"))
            return ctlab_last_error();
    R->have_data = true;
    return CTLAB_OK;
}

int ctlab_os_sart(ctlab_recon* R, float* d_x, int iters, float lambda, int nonneg, int use_graph, void* stream)
{
    clear_error();
    if (!R || !d_x || iters < 0 || !R->have_data) {
        set_error(CTLAB_ERR_ARG, "need recon with data, volume and iters >= 0");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(R->device);
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        for (int it = 0; it < iters; ++it)
            if (sart_pass(R, d_x, lambda, nonneg, s)) return ctlab_last_error();
        return CTLAB_OK;
    }
    // capture one pass on a private stream (the graph reads d_x in place), replay `iters` times
    if (!R->graph_stream && chk(hipStreamCreateWithFlags(&R->graph_stream, hipStreamNonBlocking), "stream"))
        return ctlab_last_error();
    if (!R->sart_graph || R->graph_lambda != lambda || R->graph_nonneg != nonneg || R->graph_x != d_x) {
        if (R->sart_graph) { hipGraphExecDestroy(R->sart_graph); R->sart_graph = nullptr; }
        hipGraph_t gr = nullptr;
        if (chk(hipStreamBeginCapture(R->graph_stream, hipStreamCaptureModeThreadLocal), "begin capture"))
            return ctlab_last_error();
        int rc = sart_pass(R, d_x, lambda, nonneg, R->graph_stream);
        hipError_t e = hipStreamEndCapture(R->graph_stream, &gr);
        if (rc || chk(e, "end capture")) { if (gr) hipGraphDestroy(gr); return ctlab_last_error(); }
        e = hipGraphInstantiate(&R->sart_graph, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (chk(e, "graph instantiate")) return ctlab_last_error();
        R->graph_lambda = lambda;
        R->graph_nonneg = nonneg;
        R->graph_x = d_x;
    }
    // order the private stream after the caller's stream, replay, then order back
    hipEvent_t ev;
    if (chk(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event")) return ctlab_last_error();
    hipEventRecord(ev, s);
    hipStreamWaitEvent(R->graph_stream, ev, 0);
    for (int it = 0; it < iters; ++it)
        if (chk(hipGraphLaunch(R->sart_graph, R->graph_stream), "graph launch")) { hipEventDestroy(ev); return ctlab_last_error(); }
    hipEventRecord(ev, R->graph_stream);
    hipStreamWaitEvent(s, ev, 0);
    hipEventDestroy(ev);
    return CTLAB_OK;
}

// ||A x - b||^2 in FP64 (synchronous)
double ctlab_recon_data_fidelity(ctlab_recon* R, const float* d_x, void* stream)
{
    clear_error();
    if (!R || !d_x || !R->have_data) {
        set_error(CTLAB_ERR_ARG, "need recon with data and a volume");
        return NAN;
    }
    DeviceGuard dg(R->device);
    return residual(R, d_x, (hipStream_t)stream, true);
}

// g = A^T (A y - b); returns ||A y - b||^2 (synchronous)
double ctlab_recon_gradient(ctlab_recon* R, const float* d_y, float* d_g, void* stream)
{
    clear_error();
    if (!R || !d_y || !d_g || !R->have_data) {
        set_error(CTLAB_ERR_ARG, "need recon with data, y and g");
        return NAN;
    }
    DeviceGuard dg(R->device);
    hipStream_t s = (hipStream_t)stream;
    const double f = residual(R, d_y, s, true);
    if (plan_back(R->full, R->d_prj, d_g, s)) return NAN;
    return f;
}

// largest eigenvalue of A^T A by power iteration
double ctlab_recon_lipschitz(ctlab_recon* R, int iters, void* stream)
{
    clear_error();
    if (!R || iters < 1) {
        set_error(CTLAB_ERR_ARG, "need recon and iters >= 1");
        return NAN;
    }
    DeviceGuard dg(R->device);
    hipStream_t s = (hipStream_t)stream;
    if (chk(rk_fill(R->d_vol2, 1.0f, R->nvox, s), "fill")) return NAN;
    double lam = 0.0;
    for (int it = 0; it < iters; ++it) {
        const double nv = std::sqrt(dot(R, R->d_vol2, nullptr, R->nvox, s));
        if (!(nv > 0)) return 0.0;
        if (chk(rk_axpby((float)(1.0 / nv), R->d_vol2, 0.0f, R->d_vol2, R->nvox, s), "scale")) return NAN;
        if (plan_proj(R->full, R->d_vol2, R->d_prj, s) || plan_back(R->full, R->d_prj, R->d_vol3, s)) return NAN;
        lam = dot(R, R->d_vol2, R->d_vol3, R->nvox, s);   // Rayleigh quotient, ||v|| = 1
        if (chk(hipMemcpyAsync(R->d_vol2, R->d_vol3, R->nvox * sizeof(float), hipMemcpyDeviceToDevice, s), "D2D"))
            return NAN;
    }
    return lam;
}

int ctlab_tv_prox(const float* d_b, float* d_x, float lam, int iters, int XN, int YN, int ZN, float* d_work,
                  void* stream)
{
    clear_error();
    if (!d_b || !d_x || !d_work || XN <= 0 || YN <= 0 || ZN <= 0 || iters < 0) {
        set_error(CTLAB_ERR_ARG, "bad TV prox arguments");
        return CTLAB_ERR_ARG;
    }
    const long n = (long)XN * YN * ZN;
    if (chk(rk_tv_prox(d_b, d_x, lam, iters, XN, YN, ZN, d_work, (hipStream_t)stream), "tv prox"))
        return ctlab_last_error();
    return CTLAB_OK;
}

// FISTA-TV: x <- prox_{lam/L TV}(y - A^T(Ay - b)/L); y <- FISTA momentum (clamped)
int ctlab_fista_tv(ctlab_recon* R, float* d_x, int iters, float tv_lambda, int tv_iters, float L, int nonneg,
                   double* h_obj, void* stream)
{
    clear_error();
    if (!R || !d_x || iters < 0 || !R->have_data) {
        set_error(CTLAB_ERR_ARG, "need recon with data, volume and iters >= 0");
        return CTLAB_ERR_ARG;
    }
    DeviceGuard dg(R->device);
    hipStream_t s = (hipStream_t)stream;
    if (!(L > 0)) {
        L = (float)ctlab_recon_lipschitz(R, 20, stream) * 1.01f;
        if (!(L > 0)) return ctlab_last_error() ? ctlab_last_error() : CTLAB_ERR_GEOMETRY;
    }
    if (!R->d_tv && chk(hipMalloc(&R->d_tv, 9 * R->nvox * sizeof(float)), "hipMalloc")) return ctlab_last_error();
    const long n = R->nvox;
    float* y = R->d_vol2;      // extrapolated point
    float* xold = R->d_vol3;
    float* z = R->d_vol4;
    hipMemcpyAsync(y, d_x, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(xold, d_x, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    double t = 1.0;
    for (int it = 0; it < iters; ++it) {
        const double f = ctlab_recon_gradient(R, y, R->d_vol, stream);   // d_vol = A^T(Ay - b)
        if (std::isnan(f)) return ctlab_last_error();
        if (h_obj) h_obj[it] = f;
        if (chk(rk_xpaymz(z, y, -1.0f / L, R->d_vol, nullptr, n, s), "gradient step")) return ctlab_last_error();
        if (chk(rk_tv_prox(z, d_x, tv_lambda / L, tv_iters, R->XN, R->YN, R->ZN, R->d_tv, s), "tv prox"))
            return ctlab_last_error();
        const double tn = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
        // optional non-negativity: FISTA_functor with t1 = 1 is a pure clamp
        if (nonneg && chk(rk_fista(d_x, d_x, 1.0f, 1.0f, 0.0f, 6000.0f, n, s), "clamp")) return ctlab_last_error();
        hipMemcpyAsync(y, d_x, n * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (chk(rk_fista(y, xold, (float)t, (float)tn, nonneg ? 0.0f : -2000.0f, 6000.0f, n, s), "fista"))
            return ctlab_last_error();
        hipMemcpyAsync(xold, d_x, n * sizeof(float), hipMemcpyDeviceToDevice, s);
        t = tn;
    }
    return chk(hipStreamSynchronize(s), "fista") ? ctlab_last_error() : CTLAB_OK;
}

// CG on the normal equations with the reference's beta rules (cgm = CGMethod 0..4).
int ctlab_cg(ctlab_recon* R, float* d_x, int iters, int cgm, double* h_resid, void* stream)
{
    clear_error();
    return ctlab::cg_run(R, d_x, iters, cgm, h_resid, (hipStream_t)stream, nullptr);
}

// Fletcher-Reeves CG from d_x until ||Ax - b||^2 <= f_target (at most max_iters steps).
int ctlab_cg_to_fidelity(ctlab_recon* R, float* d_x, int max_iters, double f_target, double* h_fid,
                         int* iters_done, void* stream)
{
    clear_error();
    if (!iters_done || std::isnan(f_target)) {
        set_error(CTLAB_ERR_ARG, "need iters_done and a target that is not NaN");
        return CTLAB_ERR_ARG;
    }
    ctlab::CgStop stop{f_target, h_fid, 0};
    const int rc = ctlab::cg_run(R, d_x, max_iters, 0, nullptr, (hipStream_t)stream, nullptr, &stop);
    *iters_done = stop.done;
    return rc;
}

// ---- elementwise entry points (device pointers, stream-ordered) ---------
int ctlab_sart_prj_weight(float* d_prj, const float* d_real, const float* d_rowSum, long long n, void* stream)
{
    clear_error();
    return chk(rk_sart_prj_weight(d_prj, d_real, d_rowSum, n, (hipStream_t)stream), "sart") ? ctlab_last_error() : 0;
}

int ctlab_sart_bak_weight(const float* d_vol, float* d_recon, const float* d_colSum, long long n, void* stream)
{
    clear_error();
    return chk(rk_sart_bak_weight(d_vol, d_recon, d_colSum, 1.0f, 0, 0.0f, n, (hipStream_t)stream), "sart")
               ? ctlab_last_error() : 0;
}

int ctlab_ossart_update(float* d_img, const float* d_cor, const float* d_weg, const float* d_msk, float lambda,
                        long long n, void* stream)
{
    clear_error();
    return chk(rk_ossart_update(d_img, d_cor, d_weg, d_msk, lambda, n, (hipStream_t)stream), "ossart")
               ? ctlab_last_error() : 0;
}

int ctlab_fista_momentum(float* d_cur, const float* d_las, float t1, float t2, float minV, float maxV, long long n,
                         void* stream)
{
    clear_error();
    return chk(rk_fista(d_cur, d_las, t1, t2, minV, maxV, n, (hipStream_t)stream), "fista") ? ctlab_last_error() : 0;
}

int ctlab_axpby(float a, const float* d_x, float b, float* d_y, long long n, void* stream)
{
    clear_error();
    return chk(rk_axpby(a, d_x, b, d_y, n, (hipStream_t)stream), "axpby") ? ctlab_last_error() : 0;
}

int ctlab_soft_threshold(float* d_x, float g, long long n, void* stream)
{
    clear_error();
    return chk(rk_soft_threshold(d_x, g, n, (hipStream_t)stream), "soft threshold") ? ctlab_last_error() : 0;
}

}  // extern "C"

// Scratch of ctlab_cg_beta_static: the FP64 partial sums of the beta
// reduction (two per block of its fixed grid).  One array per GPU, held until
// the process ends; every call overwrites all of it before reading it, so
// nothing needs clearing, and calls that overlap on one GPU race on it.
static constexpr int kBetaScratchDoubles = 2048;
static __device__ __align__(256) double scratch[kBetaScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;   // device address of `scratch`, per device

// device address of `scratch` on the current device (looked up once per device,
// with that device current); null with the error recorded on failure
double* scratch_on_current_device()
{
    int dev = 0;
    if (chk(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (chk(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return nullptr;
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int ctlab_cg_beta(const float* d_gk, const float* d_gk1, const float* d_d, int cases, long long n, double* h_numden,
                  void* stream)
{
    clear_error();
    if (!h_numden || cases < 0 || cases > 4) {
        set_error(CTLAB_ERR_ARG, "bad beta arguments");
        return CTLAB_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr, *red = nullptr;
    if (chk(hipMalloc(&part, sizeof(double) * red_scratch_doubles()), "hipMalloc") ||
        chk(hipMalloc(&red, sizeof(double) * 2), "hipMalloc") ||
        chk(rk_beta(d_gk, d_gk1, d_d, cases, 0, n, part, red, s), "beta") ||
        chk(hipMemcpyAsync(h_numden, red, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H") ||
        chk(hipStreamSynchronize(s), "sync")) {
        if (part) hipFree(part);
        if (red) hipFree(red);
        return ctlab_last_error();
    }
    hipFree(part);
    hipFree(red);
    return CTLAB_OK;
}

// ctlab_cg_beta without any allocation: the partial sums live in the
// file-scope `scratch` array and the two sums go to the caller's device
// buffer d_numden, stream-ordered (no synchronisation), so the call can be
// captured in a HIP graph.  The scratch need is fixed (two doubles per block
// of the reduction grid), whatever n.
int ctlab_cg_beta_static(const float* d_gk, const float* d_gk1, const float* d_d, int cases, long long n,
                         double* d_numden, void* stream)
{
    clear_error();
    if (!d_gk || !d_gk1 || !d_d || !d_numden || cases < 0 || cases > 4 || n < 0) {
        set_error(CTLAB_ERR_ARG, "bad beta arguments");
        return CTLAB_ERR_ARG;
    }
    if (red_scratch_doubles() > kBetaScratchDoubles) {
        set_error(CTLAB_ERR_ARG, "reduction needs %d doubles of scratch, the static array holds %d",
                  red_scratch_doubles(), kBetaScratchDoubles);
        return CTLAB_ERR_ARG;
    }
    double* part = scratch_on_current_device();
    if (!part) return ctlab_last_error();
    return chk(rk_beta(d_gk, d_gk1, d_d, cases, 0, n, part, d_numden, (hipStream_t)stream), "beta")
               ? ctlab_last_error()
               : CTLAB_OK;
}

}  // extern "C"

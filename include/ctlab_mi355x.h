/*
 * ctlab_mi355x.h -- C ABI of the MI355X-native branchless distance-driven
 * projector / backprojector library (libctlab_mi355x.so, gfx950).
 *
 * Part 1 replaces CTLAB's own extern "C" entry points symbol for symbol.
 * Part 2 is this library's device-resident API (no counterpart in CTLAB):
 * the reference entry points take host pointers and copy on every call; the
 * plan API keeps geometry tables and workspaces on the GPU so reconstruction
 * loops can chain projector, backprojector and update kernels on one stream.
 *
 * Data layout (CTLAB leaves the projection layout unstated; fixed here):
 *   volume     vol[(y*XN + x)*ZN + z]   ((Z,X,Y) order, src/utilities.cuh:441-454)
 *   projection prj[(p*DNU + u)*DNV + v]
 *   mask       mask[y*XN + x]           (voxel used / written iff mask != 0; NULL = all)
 *   multi-slice (DD2_multiGPU): SLN plays the role of ZN and DNV.
 * Errors never exit or throw across the ABI: each call records a status read by
 * ctlab_last_error() (per host thread); on error the output is left untouched.
 */
#ifndef CTLAB_MI355X_H_
#define CTLAB_MI355X_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned char byte;

/* ---- Part 1: CTLAB entry points (prjMode / method 0 and 2 only) ---------- */

/* src/DD3_GPU_Proj.h:41-44.  Cone-beam branchless DD projection, prjMode 0;
 * other modes set CTLAB_ERR_MODE.  Host pointers; hprj is overwritten. */
void DD3Proj_gpu(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                 float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos,
                 int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz,
                 byte* mask, int gpunum, int prjMode);

/* DD3Proj_gpu / DD3Back_gpu / DD2_multiGPU keep, per geometry and device, the
 * plan (per-view tables, integral-image workspace) and device buffers of their
 * last calls: a repeated call with the same geometry runs only H2D, the
 * operator and D2H.  ctlab_plan_cache_* (this library's own) bound or free
 * that memory.  A failed call also prints its message to stderr, since these
 * entry points return void (CTLAB_QUIET=1 silences it). */
/* src/DD3_GPU_Back.h:40-44.  Exact adjoint of DD3Proj_gpu, prjMode 0;
 * `squared` is accepted and ignored (the reference calls it a fake parameter,
 * src/DD3_GPU_Back.h:38).  hvol is overwritten. */
void DD3Back_gpu(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                 float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos,
                 int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz,
                 byte* mask, int gpunum, int squared, int prjMode);

/* src/multiSlices_ker.cuh:35-38.  method 0 = branchless projection, 2 =
 * branchless backprojection (exact adjoint); 1 / 3 set CTLAB_ERR_MODE.  All
 * SLN slices run on the current device (startIdx must start at 0 and be
 * non-decreasing; gpuNum >= 1 is accepted, one GPU is used). */
void DD2_multiGPU(float* hvol, float* hprj, const int method, const float x0, const float y0,
                  float* xds, float* yds, const int DNU, const int SLN, const float imgXCenter,
                  const float imgYCenter, const int XN, const int YN, const float dx, float* hangs,
                  int PN, byte* mask, int* startIdx, const int gpuNum);

/* Multi-GPU split entry points.  The reference splits the views (projector,
 * startPN) or the volume slices (backprojector, startVOL) over GPUs
 * 0..gpuNum-1; every part is the same operator, so the union equals one
 * single-GPU call, which is what runs here, on the current device, after the
 * split is checked (index 0 first, non-decreasing, within PN / ZN).  Mode /
 * methodId other than 0 set CTLAB_ERR_MODE.  The Helical_3GPU / _4GPU forms
 * take `int (&)[3]` / `int (&)[4]` in the reference, a pointer in this ABI. */
/* src/DD3_GPU_Proj.h:138-141 */
void DD3Proj_multiGPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                      float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos,
                      int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz,
                      byte* mask, int prjMode, int* startPN, int gpuNum);
/* src/DD3_GPU_Back.h:138-142 */
void DD3Back_multiGPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds, float* zds,
                      float imgXCenter, float imgYCenter, float imgZCenter, float* hangs, float* hzPos,
                      int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx, float dz,
                      byte* mask, int methodId, int* startVOL, int gpuNum);
/* src/DD3_GPU_Proj.h:73-76, 105-108 */
void DD3ProjHelical_3GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds,
                         float* zds, float imgXCenter, float imgYCenter, float imgZCenter, float* hangs,
                         float* hzPos, int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx,
                         float dz, byte* mask, int methodId, int* startPN);
void DD3ProjHelical_4GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds,
                         float* zds, float imgXCenter, float imgYCenter, float imgZCenter, float* hangs,
                         float* hzPos, int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx,
                         float dz, byte* mask, int methodId, int* startPN);
/* src/DD3_GPU_Back.h:73-76, 105-108 */
void DD3BackHelical_3GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds,
                         float* zds, float imgXCenter, float imgYCenter, float imgZCenter, float* hangs,
                         float* hzPos, int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx,
                         float dz, byte* mask, int methodId, int* startVOL);
void DD3BackHelical_4GPU(float x0, float y0, float z0, int DNU, int DNV, float* xds, float* yds,
                         float* zds, float imgXCenter, float imgYCenter, float imgZCenter, float* hangs,
                         float* hzPos, int PN, int XN, int YN, int ZN, float* hvol, float* hprj, float dx,
                         float dz, byte* mask, int methodId, int* startVOL);

/* ---- status --------------------------------------------------------------- */
enum {
    CTLAB_OK = 0,
    CTLAB_ERR_ARG = 1,       /* invalid argument (sizes, NULL, non-finite geometry) */
    CTLAB_ERR_MODE = 2,      /* unsupported prjMode / method */
    CTLAB_ERR_GEOMETRY = 3,  /* unsupported geometry (source inside volume, uneven rows, ...) */
    CTLAB_ERR_HIP = 4,       /* HIP runtime error */
    CTLAB_ERR_NOMEM = 5,     /* device allocation failed */
    CTLAB_ERR_IO = 6         /* file could not be opened, read or written in full */
};
int ctlab_last_error(void);
const char* ctlab_error_string(int code);
const char* ctlab_last_error_message(void);

/* ---- Part 2: device-resident plan API (this library's own) --------------- */
typedef struct ctlab_plan ctlab_plan;

/* Cone-beam plan from the DD3Proj_gpu geometry arguments (host arrays). */
ctlab_plan* ctlab_plan_create(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                              const float* yds, const float* zds, float imgXCenter,
                              float imgYCenter, float imgZCenter, const float* hangs,
                              const float* hzPos, int PN, int XN, int YN, int ZN, float dx,
                              float dz, const byte* mask, int gpunum);
/* Multi-slice fan-beam plan from the DD2_multiGPU geometry arguments. */
ctlab_plan* ctlab_plan_create_fan(float x0, float y0, const float* xds, const float* yds, int DNU,
                                  int SLN, float imgXCenter, float imgYCenter, int XN, int YN,
                                  float dx, const float* hangs, int PN, const byte* mask,
                                  int gpunum);
void ctlab_plan_destroy(ctlab_plan* plan);
/* precision of the integral images: 64 (default) or 32 */
int ctlab_plan_set_sat_bits(ctlab_plan* plan, int bits);
/* d_vol, d_prj: device pointers in the layouts above; stream: hipStream_t or NULL */
int ctlab_proj(ctlab_plan* plan, const float* d_vol, float* d_prj, void* stream);
int ctlab_back(ctlab_plan* plan, const float* d_prj, float* d_vol, void* stream);
/* tile shapes: detector columns per projector wavefront (nu) and voxel rows
 * per backprojector wavefront (nm), each 8, 12 or 16 (default 8 / 16) */
int ctlab_plan_set_tiles(ctlab_plan* plan, int nu, int nm);
int ctlab_plan_get_tiles(const ctlab_plan* plan, int* nu, int* nm);
/* autotuner: times every tile variant of both operators on these device
 * buffers (contents overwritten) and keeps the fastest; ms_out (6 floats or
 * NULL) receives proj 8/12/16 then back 8/12/16 milliseconds */
int ctlab_plan_autotune(ctlab_plan* plan, float* d_vol, float* d_prj, float* ms_out);
/* geometry queries */
int ctlab_plan_views(const ctlab_plan* plan, int* n_yslab, int* n_xslab);
long long ctlab_plan_workspace_bytes(const ctlab_plan* plan);
/* plan cache of the host-pointer entry points: at most `entries` geometries
 * (default 16; 0 frees and disables it); returns the previous limit */
int ctlab_plan_cache_set_limit(int entries);
void ctlab_plan_cache_clear(void);
int ctlab_plan_cache_size(void);

/* ---- reconstruction on the device (this library's own) ------------------ */
typedef struct ctlab_recon ctlab_recon;

/* OS-SART / FISTA-TV / CG state over a plan's geometry: osNum interleaved view
 * subsets (view p belongs to subset p % osNum), their row sums A_s*1 and column
 * sums A_s^T*1 (src/DD3_GPU_recon.h:21-28), and work buffers, all in HBM. */
ctlab_recon* ctlab_recon_create(const ctlab_plan* plan, int osNum);
void ctlab_recon_destroy(ctlab_recon* recon);
/* tiles of every operator the reconstruction runs (see ctlab_plan_set_tiles) */
int ctlab_recon_set_tiles(ctlab_recon* recon, int nu, int nm);
/* measured projections (device pointer, plan layout, original view order) */
int ctlab_recon_set_data(ctlab_recon* recon, const float* d_b, void* stream);
/* OS-SART (src/DD3_GPU_recon.h:60-65, weighting src/SARTWeighting.h:22-91):
 * x += lambda * A_s^T((b_s - A_s x)/rowSum_s)/colSum_s per subset; nonneg clamps
 * at 0; use_graph captures one pass over all subsets in a HIP graph and replays it. */
int ctlab_os_sart(ctlab_recon* recon, float* d_x, int iters, float lambda, int nonneg, int use_graph,
                  void* stream);
/* FISTA with the FGP TV prox: x <- prox_{tv_lambda/L TV}(y - A^T(Ay-b)/L), y by the
 * FISTA momentum (src/FastIterativeShrinkageAlgorithm.h:17-42); L <= 0 estimates
 * ||A||^2 by power iteration; h_obj (iters, or NULL) receives ||Ay-b||^2. */
int ctlab_fista_tv(ctlab_recon* recon, float* d_x, int iters, float tv_lambda, int tv_iters, float L,
                   int nonneg, double* h_obj, void* stream);
/* CG on the normal equations, beta rule cgm = CG_FR..CG_DY (src/DD3_GPU_recon.h:14-19,
 * src/utilities.hpp:1228-1268); h_resid (iters, or NULL) receives ||A^T(b-Ax)||^2. */
int ctlab_cg(ctlab_recon* recon, float* d_x, int iters, int cgm, double* h_resid, void* stream);
/* No counterpart in the reference: the same CG (Fletcher-Reeves) from d_x, run
 * until ||Ax - b||^2 <= f_target or max_iters steps.  The value is carried along
 * exactly (each step lowers it by 2 alpha <A^T(b - Ax), d> - alpha^2 ||Ad||^2, no
 * extra projection); h_fid (max_iters, or NULL) receives it after each step and
 * *iters_done the steps taken.  For callers that compare candidates at equal
 * data residual (the learned step's data-consistency stage). */
int ctlab_cg_to_fidelity(ctlab_recon* recon, float* d_x, int max_iters, double f_target, double* h_fid,
                         int* iters_done, void* stream);
/* ||Ax - b||^2 (FP64) */
double ctlab_recon_data_fidelity(ctlab_recon* recon, const float* d_x, void* stream);
/* d_g = A^T(Ay - b); returns ||Ay - b||^2 */
double ctlab_recon_gradient(ctlab_recon* recon, const float* d_y, float* d_g, void* stream);
/* largest eigenvalue of A^T A (power iteration) */
double ctlab_recon_lipschitz(ctlab_recon* recon, int iters, void* stream);

/* ---- CTLAB's host-level functions around the projector -------------------
 * C-ABI forms of C++ functions of the reference (thrust::host_vector and
 * std::string arguments become pointers and sizes); include/ctlab_mi355x.hpp
 * restores the reference's own names and signatures for C++ callers. */

/* src/utilities.cuh:478-499 getVolZIdxPair: slices [*objStart, *objEnd) that
 * views [prjStart, prjEnd) of a helical scan can reach (zPos ascending; row
 * centre index detCntIdxV, row pitch detStpZ at the detector, volume centre
 * index objCntIdxZ), clamped to [0, ZN]. */
int ctlab_getVolZIdxPair(const float* zPos, int PN, int prjStart, int prjEnd, float detCntIdxV, float detStpZ,
                         int DNV, float objCntIdxZ, float dz, int ZN, int* objStart, int* objEnd);
/* src/utilities.cuh:505-532 getPrjIdxPair: views [*prjStart, *prjEnd) whose
 * detector can reach slices [objStart, objEnd), clamped to [0, PN]. */
int ctlab_getPrjIdxPair(const float* zPos, int PN, int objStart, int objEnd, float objCntIdxZ, float dz, int ZN,
                        float detCntIdxV, float detStpZ, int DNV, int* prjStart, int* prjEnd);

/* src/readwrite.hpp:15-113 readData_host / writeData_host /
 * writeFileFromDeviceToDisk: n raw native-endian elements of elemBytes (4:
 * float, 8: double); host or device memory.  A short file is CTLAB_ERR_IO and
 * leaves the buffer untouched. */
int ctlab_read_raw(const char* fileName, void* h_data, long long n, int elemBytes);
int ctlab_write_raw(const char* fileName, const void* h_data, long long n, int elemBytes);
int ctlab_read_raw_device(const char* fileName, void* d_data, long long n, int elemBytes);
int ctlab_write_raw_device(const char* fileName, const void* d_data, long long n, int elemBytes);

/* src/DD3_GPU_recon.h:21-28 evalProjBack: A*1 (forwardProjection != 0, into
 * a projection-sized d_out) or A^T*1 (volume-sized d_out) over a plan. */
int ctlab_evalProjBack(ctlab_plan* plan, int forwardProjection, int forwardMethod, int backwardMethod, float* d_out,
                       void* stream);

/* src/DD3_GPU_recon.h:37-65 OS_SART and :67-103 CG, same arguments in the
 * same order.  Geometry: source at (0, sid, 0) at view 0, equiangular arc
 * detector centred on the source (cell u at angle (u - (DNU-1)/2 - col_offset)
 * * col_size / sdd, row v at (v - (DNV-1)/2 - row_offset) * row_size), views
 * hangs / hzPos.  initImg NULL starts from 0.  OS-SART: lambda 1, osNum
 * interleaved subsets; CG: all views (osNum ignored), beta rule cgm.  Runs on
 * the current device; reconImg receives the result, which is also written
 * to volumeName (raw float, (Z,X,Y) order) unless volumeName is NULL or "";
 * outputMedRes writes iteration k to volumeName_<k>.  Only method 0 (branchless)
 * is built: other forward / backward methods give CTLAB_ERR_MODE. */
/* the detector cell centres (view 0) that ctlab_OS_SART / ctlab_CG build
 * from their scan arguments: xds, yds (DNU), zds (DNV) */
int ctlab_arc_detector(float sid, float sdd, int DNU, int DNV, float col_size, float row_size, float col_offset,
                       float row_offset, float* xds, float* yds, float* zds);
int ctlab_OS_SART(float* reconImg, const float* initImg, const float* hprj, const float* hangs, const float* hzPos,
                  const byte* mask, const char* volumeName, int osNum, int iterNum, int outputMedRes, float sid,
                  float sdd, int DNU, int DNV, int PN, float imgXCenter, float imgYCenter, float imgZCenter, int XN,
                  int YN, int ZN, float dx, float dz, float col_size, float row_size, float col_offset,
                  float row_offset, int forwardMethod, int backwardMethod);
int ctlab_CG(float* reconImg, const float* initImg, const float* hprj, const float* hangs, const float* hzPos,
             const byte* mask, const char* volumeName, int osNum, int iterNum, int outputMedRes, float sid, float sdd,
             int DNU, int DNV, int PN, float imgXCenter, float imgYCenter, float imgZCenter, int XN, int YN, int ZN,
             float dx, float dz, float col_size, float row_size, float col_offset, float row_offset,
             int forwardMethod, int backwardMethod, int cgm);

/* ---- batched OS-SART (this library's own; the reference runs one problem per
 * call, src/DD3_GPU_recon.h:60-65) -----------------------------------------
 * n reconstructions (each with data set, one device, one SAT precision; any
 * volume, view and detector sizes and subset counts) advance together: every
 * phase of a subset step is one persistent launch over all problems, and one
 * iteration is captured in a HIP graph.  Results equal n separate
 * ctlab_os_sart calls bit for bit.  The batch borrows the reconstructions'
 * buffers: do not run them on their own while the batch is in use, and
 * destroy the batch first. */
typedef struct ctlab_batch ctlab_batch;
ctlab_batch* ctlab_batch_create(ctlab_recon* const* recons, int n);
void ctlab_batch_destroy(ctlab_batch* batch);
int ctlab_batch_size(const ctlab_batch* batch);
/* tiles of the batched projector (nu) and backprojector (nm): 8, 12 or 16 */
int ctlab_batch_set_tiles(ctlab_batch* batch, int nu, int nm);
int ctlab_batch_get_tiles(const ctlab_batch* batch, int* nu, int* nm);
/* iters OS-SART iterations on d_x[0..n-1] (device volumes, updated in place) */
int ctlab_batch_os_sart(ctlab_batch* batch, float* const* d_x, int iters, float lambda, int nonneg,
                        int use_graph, void* stream);
/* runs one iteration per tile size on d_x (updated), keeps the fastest
 * projector and backprojector tiles; ms_out (6 floats or NULL): projection
 * phases for nu = 8/12/16, then backprojection phases for nm = 8/12/16 */
int ctlab_batch_autotune(ctlab_batch* batch, float* const* d_x, float lambda, float* ms_out);

/* elementwise / reduction kernels on device buffers (stream-ordered):
 * SART_Weighting::Proj / Back (src/SARTWeighting.h:97-136), ossart_update_functor
 * (src/utilities.hpp:1202-1223), FISTA_functor (src/FastIterativeShrinkageAlgorithm.h:
 * 17-42), _softThreshold_functor (src/utilities.hpp:1097-1117), calculateBetaForCG
 * (src/utilities.hpp:1611-1621; h_numden receives the two FP64 sums). */
int ctlab_sart_prj_weight(float* d_prj, const float* d_real, const float* d_rowSum, long long n, void* stream);
int ctlab_sart_bak_weight(const float* d_vol, float* d_recon, const float* d_colSum, long long n, void* stream);
int ctlab_ossart_update(float* d_img, const float* d_cor, const float* d_weg, const float* d_msk, float lambda,
                        long long n, void* stream);
int ctlab_fista_momentum(float* d_cur, const float* d_las, float t1, float t2, float minV, float maxV,
                         long long n, void* stream);
int ctlab_soft_threshold(float* d_x, float g, long long n, void* stream);
/* d_y = a*d_x + b*d_y (the CG_update1/2/3 form, src/utilities.hpp:1121-1156) */
int ctlab_axpby(float a, const float* d_x, float b, float* d_y, long long n, void* stream);
int ctlab_cg_beta(const float* d_gk, const float* d_gk1, const float* d_d, int cases, long long n,
                  double* h_numden, void* stream);
/* No counterpart in the reference.  ctlab_cg_beta for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): no allocation, no synchronisation; the two
 * FP64 sums go to d_numden (2 doubles, device memory), stream-ordered.  Its
 * scratch (the reduction's partial sums, 16 KiB, fixed whatever n) is one
 * static array per GPU shared by every call on that GPU: calls that overlap on
 * one GPU race. */
int ctlab_cg_beta_static(const float* d_gk, const float* d_gk1, const float* d_d, int cases, long long n,
                         double* d_numden, void* stream);
/* isotropic TV prox (FGP, iters steps) on a (Z,X,Y) volume; d_work: 9*XN*YN*ZN floats */
int ctlab_tv_prox(const float* d_b, float* d_x, float lam, int iters, int XN, int YN, int ZN, float* d_work,
                  void* stream);

/* ---- learned proximal CNN (no counterpart in CTLAB) ----------------------
 * 5 layers of 3x3x3 convolution 1 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after
 * layers 1-4, output x + net(x); BF16 weights / activations on MFMA with FP32
 * accumulation.  Parameters: one flat FP32 vector W1 b1 W2 b2 ... W5 b5 with
 * W_l[27][cout][cin], tap t = (dy+1)*9 + (dx+1)*3 + (dz+1).  An object is sized
 * for one volume shape; training needs ZN % 8 == 0. */
typedef struct ctlab_cnn ctlab_cnn;
int ctlab_cnn_num_params(void);
ctlab_cnn* ctlab_cnn_create(int XN, int YN, int ZN, int training, int gpunum);
void ctlab_cnn_destroy(ctlab_cnn* cnn);
int ctlab_cnn_set_params(ctlab_cnn* cnn, const float* h_params);
int ctlab_cnn_get_params(ctlab_cnn* cnn, float* h_params);
/* d_out = d_x + net(d_x) (device pointers, volume layout) */
int ctlab_cnn_forward(ctlab_cnn* cnn, const float* d_x, float* d_out, void* stream);
/* forward + backward of 0.5*mean((x + net(x) - target)^2); returns the loss,
 * keeps the gradient on the device (copied to h_grad when non-NULL) */
double ctlab_cnn_loss_grad(ctlab_cnn* cnn, const float* d_x, const float* d_target, float* d_out, float* h_grad,
                           void* stream);
/* one Adam step with the gradient of the last ctlab_cnn_loss_grad */
int ctlab_cnn_adam_step(ctlab_cnn* cnn, float lr, float beta1, float beta2, float eps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CTLAB_MI355X_H_ */

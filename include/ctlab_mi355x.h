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

/* ---- status --------------------------------------------------------------- */
enum {
    CTLAB_OK = 0,
    CTLAB_ERR_ARG = 1,       /* invalid argument (sizes, NULL, non-finite geometry) */
    CTLAB_ERR_MODE = 2,      /* unsupported prjMode / method */
    CTLAB_ERR_GEOMETRY = 3,  /* unsupported geometry (source inside volume, ZN or DNV > 64, ...) */
    CTLAB_ERR_HIP = 4,       /* HIP runtime error */
    CTLAB_ERR_NOMEM = 5      /* device allocation failed */
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
/* geometry queries */
int ctlab_plan_views(const ctlab_plan* plan, int* n_yslab, int* n_xslab);
long long ctlab_plan_workspace_bytes(const ctlab_plan* plan);

#ifdef __cplusplus
}
#endif
#endif /* CTLAB_MI355X_H_ */

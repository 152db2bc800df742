// Host-level entry points around the projector that CTLAB's users call:
//   * the helical view / slice range helpers getVolZIdxPair / getPrjIdxPair
//     (src/utilities.cuh:476-532);
//   * raw file input / output (src/readwrite.hpp:15-113; the bodies are not in
//     the reference tree: raw native-endian arrays of n elements, host or device
//     memory);
//   * evalProjBack (src/DD3_GPU_recon.h:21-28) and the OS_SART / CG drivers with
//     the argument lists of src/DD3_GPU_recon.h:37-103 (bodies absent too):
//     geometry from sid, sdd, detector size and offsets; device-resident
//     iterations; the result written to VolumeName in (Z,X,Y) float order and,
//     with outputMedRes, every iteration to VolumeName_<k>.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctlab_mi355x.h"
#include "plan.hpp"
#include "recon.hpp"
#include "status.hpp"

namespace ctlab {
hipError_t rk_fill(float* x, float v, long n, hipStream_t s);
}
using namespace ctlab;

namespace {

int elem_ok(int elemBytes)
{
    if (elemBytes == 4 || elemBytes == 8) return 1;
    set_error(CTLAB_ERR_ARG, "element size must be 4 (float) or 8 (double), got %d", elemBytes);
    return 0;
}

int read_file(const char* name, void* dst, size_t bytes)
{
    FILE* f = std::fopen(name, "rb");
    if (!f) {
        set_error(CTLAB_ERR_IO, "cannot open %s for reading", name);
        return CTLAB_ERR_IO;
    }
    const size_t got = bytes ? std::fread(dst, 1, bytes, f) : 0;
    std::fclose(f);
    if (got != bytes) {
        set_error(CTLAB_ERR_IO, "%s holds %zu of the %zu bytes requested", name, got, bytes);
        return CTLAB_ERR_IO;
    }
    return CTLAB_OK;
}

int write_file(const char* name, const void* src, size_t bytes)
{
    FILE* f = std::fopen(name, "wb");
    if (!f) {
        set_error(CTLAB_ERR_IO, "cannot open %s for writing", name);
        return CTLAB_ERR_IO;
    }
    const size_t put = bytes ? std::fwrite(src, 1, bytes, f) : 0;
    const int closed = std::fclose(f);
    if (put != bytes || closed != 0) {
        set_error(CTLAB_ERR_IO, "writing %s failed after %zu of %zu bytes", name, put, bytes);
        return CTLAB_ERR_IO;
    }
    return CTLAB_OK;
}

// Geometry of the OS_SART / CG drivers: source at (0, sid, 0) at view 0 and an
// equiangular arc detector centred on the source, cell u at angle
// (u - (DNU-1)/2 - col_offset) * col_size / sdd, rows (v - (DNV-1)/2 - row_offset)
// * row_size (the same construction as ctlab_amd/geometry.py:arc_detector).
void arc_detector(float sid, float sdd, int DNU, int DNV, float col_size, float row_size, float col_offset,
                  float row_offset, std::vector<float>& xds, std::vector<float>& yds, std::vector<float>& zds)
{
    xds.resize(DNU);
    yds.resize(DNU);
    zds.resize(DNV);
    const double dgam = (double)col_size / (double)sdd;
    for (int u = 0; u < DNU; ++u) {
        const double g = ((double)u - ((DNU - 1) / 2.0 + (double)col_offset)) * dgam;
        xds[u] = (float)((double)sdd * std::sin(g));
        yds[u] = (float)((double)sid - (double)sdd * std::cos(g));
    }
    for (int v = 0; v < DNV; ++v)
        zds[v] = (float)(((double)v - ((DNV - 1) / 2.0 + (double)row_offset)) * (double)row_size);
}

// Everything the drivers hold on the device; released by the destructor.
struct Driver {
    ctlab_plan* plan = nullptr;
    ctlab_recon* recon = nullptr;
    float* d_b = nullptr;
    float* d_x = nullptr;
    long long nvox = 0, nprj = 0;
    ~Driver()
    {
        if (d_b) hipFree(d_b);
        if (d_x) hipFree(d_x);
        if (recon) ctlab_recon_destroy(recon);
        if (plan) ctlab_plan_destroy(plan);
    }
};

// Plan, reconstruction state and device copies of data and initial volume.
int driver_setup(Driver& D, const float* initImg, const float* hprj, const float* hangs, const float* hzPos,
                 const byte* mask, int osNum, float sid, float sdd, int DNU, int DNV, int PN, float imgXCenter,
                 float imgYCenter, float imgZCenter, int XN, int YN, int ZN, float dx, float dz, float col_size,
                 float row_size, float col_offset, float row_offset, int forwardMethod, int backwardMethod)
{
    if (forwardMethod != 0 || backwardMethod != 0) {
        set_error(CTLAB_ERR_MODE, "only the branchless models (forward / backward method 0) are built, got %d / %d",
                  forwardMethod, backwardMethod);
        return CTLAB_ERR_MODE;
    }
    if (!hprj || !hangs || DNU <= 0 || DNV <= 0 || PN <= 0 || XN <= 0 || YN <= 0 || ZN <= 0 ||
        !(sdd > sid) || !(sid > 0) || !(col_size > 0) || !(row_size > 0)) {
        set_error(CTLAB_ERR_ARG, "need projections, view angles, positive sizes and sdd > sid > 0");
        return CTLAB_ERR_ARG;
    }
    std::vector<float> xds, yds, zds;
    arc_detector(sid, sdd, DNU, DNV, col_size, row_size, col_offset, row_offset, xds, yds, zds);
    std::vector<float> zpos(hzPos ? hzPos : nullptr, hzPos ? hzPos + PN : nullptr);
    if (!hzPos) zpos.assign(PN, 0.0f);
    int dev = 0;
    if (hip_check(hipGetDevice(&dev), "hipGetDevice")) return ctlab_last_error();
    D.plan = ctlab_plan_create(0.0f, sid, 0.0f, DNU, DNV, xds.data(), yds.data(), zds.data(), imgXCenter,
                               imgYCenter, imgZCenter, hangs, zpos.data(), PN, XN, YN, ZN, dx, dz, mask, dev);
    if (!D.plan) return ctlab_last_error();
    D.recon = ctlab_recon_create(D.plan, osNum);
    if (!D.recon) return ctlab_last_error();
    D.nvox = (long long)XN * YN * ZN;
    D.nprj = (long long)PN * DNU * DNV;
    if (hip_check(hipMalloc(&D.d_b, D.nprj * sizeof(float)), "hipMalloc(projections)") ||
        hip_check(hipMalloc(&D.d_x, D.nvox * sizeof(float)), "hipMalloc(volume)") ||
        hip_check(hipMemcpy(D.d_b, hprj, D.nprj * sizeof(float), hipMemcpyHostToDevice), "H2D") ||
        (initImg ? hip_check(hipMemcpy(D.d_x, initImg, D.nvox * sizeof(float), hipMemcpyHostToDevice), "H2D")
                 : hip_check(hipMemset(D.d_x, 0, D.nvox * sizeof(float)), "memset")))
        return ctlab_last_error();
    return ctlab_recon_set_data(D.recon, D.d_b, nullptr);
}

// volume (device) -> host buffer and, when name is set, to a raw file
int save_volume(const Driver& D, std::vector<float>& h, const char* name)
{
    h.resize(D.nvox);
    if (hip_check(hipMemcpy(h.data(), D.d_x, D.nvox * sizeof(float), hipMemcpyDeviceToHost), "D2H"))
        return ctlab_last_error();
    if (name && name[0]) return write_file(name, h.data(), h.size() * sizeof(float));
    return CTLAB_OK;
}

std::string iteration_name(const char* volumeName, int k) { return std::string(volumeName) + "_" + std::to_string(k); }

}  // namespace

extern "C" {

int ctlab_arc_detector(float sid, float sdd, int DNU, int DNV, float col_size, float row_size, float col_offset,
                       float row_offset, float* xds, float* yds, float* zds)
{
    clear_error();
    if (!xds || !yds || !zds || DNU <= 0 || DNV <= 0 || !(sdd > 0)) {
        set_error(CTLAB_ERR_ARG, "need output arrays, DNU, DNV > 0 and sdd > 0");
        return CTLAB_ERR_ARG;
    }
    std::vector<float> x, y, z;
    arc_detector(sid, sdd, DNU, DNV, col_size, row_size, col_offset, row_offset, x, y, z);
    std::copy(x.begin(), x.end(), xds);
    std::copy(y.begin(), y.end(), yds);
    std::copy(z.begin(), z.end(), zds);
    return CTLAB_OK;
}

int ctlab_getVolZIdxPair(const float* zPos, int PN, int prjStart, int prjEnd, float detCntIdxV, float detStpZ,
                         int DNV, float objCntIdxZ, float dz, int ZN, int* objStart, int* objEnd)
{
    clear_error();
    if (!zPos || !objStart || !objEnd || PN <= 0 || prjStart < 0 || prjEnd > PN || prjStart >= prjEnd ||
        DNV <= 0 || ZN <= 0 || !(dz > 0)) {
        set_error(CTLAB_ERR_ARG, "need zPos, 0 <= prjStart < prjEnd <= PN, DNV, ZN > 0 and dz > 0");
        return CTLAB_ERR_ARG;
    }
    // detector extent below and above the source plane, at the detector; the
    // reference's (detCntIdxV + 0.5) * detStpZ is a double expression, rounded
    // to float once when stored
    const float below = (float)(((double)detCntIdxV + 0.5) * (double)detStpZ);
    const float above = (float)DNV * detStpZ - below;
    const float lo = zPos[prjStart] - below;
    const float hi = zPos[prjEnd - 1] + above;
    const int s = (int)std::floor(lo / dz + objCntIdxZ - 1.0f);
    const int e = (int)std::ceil(hi / dz + objCntIdxZ + 1.0f) + 1;
    *objStart = std::min(std::max(s, 0), ZN);
    *objEnd = std::min(std::max(e, 0), ZN);
    return CTLAB_OK;
}

int ctlab_getPrjIdxPair(const float* zPos, int PN, int objStart, int objEnd, float objCntIdxZ, float dz, int ZN,
                        float detCntIdxV, float detStpZ, int DNV, int* prjStart, int* prjEnd)
{
    clear_error();
    if (!zPos || !prjStart || !prjEnd || PN <= 0 || objStart < 0 || objEnd > ZN || objStart >= objEnd ||
        DNV <= 0 || !(dz > 0)) {
        set_error(CTLAB_ERR_ARG, "need zPos, 0 <= objStart < objEnd <= ZN, DNV > 0 and dz > 0");
        return CTLAB_ERR_ARG;
    }
    for (int p = 1; p < PN; ++p)
        if (zPos[p] < zPos[p - 1]) {
            set_error(CTLAB_ERR_ARG, "zPos must be sorted ascending (view %d)", p);
            return CTLAB_ERR_ARG;
        }
    // z extent of the sub-volume and the source positions whose detector can
    // reach it; the terms with 0.5 are double expressions in the reference,
    // rounded to float once when stored
    const float lowV = (float)(((double)((float)objStart - objCntIdxZ) - 0.5) * (double)dz);
    const float highV = lowV + (float)(objEnd - objStart) * dz;
    const float below = (float)(((double)detCntIdxV + 0.5) * (double)detStpZ);
    const float above = (float)DNV * detStpZ - below;
    const float srcLo = lowV - above;
    const float srcHi = highV + below;
    const int s = (int)(std::upper_bound(zPos, zPos + PN, srcLo) - zPos) - 1;
    const int e = (int)(std::upper_bound(zPos, zPos + PN, srcHi) - zPos) + 2;
    *prjStart = std::min(std::max(s, 0), PN);
    *prjEnd = std::min(std::max(e, 0), PN);
    return CTLAB_OK;
}

int ctlab_read_raw(const char* fileName, void* h_data, long long n, int elemBytes)
{
    clear_error();
    if (!fileName || (!h_data && n > 0) || n < 0 || !elem_ok(elemBytes)) {
        if (!ctlab_last_error()) set_error(CTLAB_ERR_ARG, "need a file name, a buffer and n >= 0");
        return ctlab_last_error();
    }
    std::vector<unsigned char> tmp((size_t)n * elemBytes);
    if (read_file(fileName, tmp.data(), tmp.size())) return ctlab_last_error();
    if (n) std::memcpy(h_data, tmp.data(), tmp.size());
    return CTLAB_OK;
}

int ctlab_write_raw(const char* fileName, const void* h_data, long long n, int elemBytes)
{
    clear_error();
    if (!fileName || (!h_data && n > 0) || n < 0 || !elem_ok(elemBytes)) {
        if (!ctlab_last_error()) set_error(CTLAB_ERR_ARG, "need a file name, a buffer and n >= 0");
        return ctlab_last_error();
    }
    return write_file(fileName, h_data, (size_t)n * elemBytes);
}

int ctlab_read_raw_device(const char* fileName, void* d_data, long long n, int elemBytes)
{
    clear_error();
    if (!fileName || (!d_data && n > 0) || n < 0 || !elem_ok(elemBytes)) {
        if (!ctlab_last_error()) set_error(CTLAB_ERR_ARG, "need a file name, a buffer and n >= 0");
        return ctlab_last_error();
    }
    std::vector<unsigned char> tmp((size_t)n * elemBytes);
    if (read_file(fileName, tmp.data(), tmp.size())) return ctlab_last_error();
    if (n && hip_check(hipMemcpy(d_data, tmp.data(), tmp.size(), hipMemcpyHostToDevice), "H2D"))
        return ctlab_last_error();
    return CTLAB_OK;
}

int ctlab_write_raw_device(const char* fileName, const void* d_data, long long n, int elemBytes)
{
    clear_error();
    if (!fileName || (!d_data && n > 0) || n < 0 || !elem_ok(elemBytes)) {
        if (!ctlab_last_error()) set_error(CTLAB_ERR_ARG, "need a file name, a buffer and n >= 0");
        return ctlab_last_error();
    }
    std::vector<unsigned char> tmp((size_t)n * elemBytes);
    if (n && hip_check(hipMemcpy(tmp.data(), d_data, tmp.size(), hipMemcpyDeviceToHost), "D2H"))
        return ctlab_last_error();
    return write_file(fileName, tmp.data(), tmp.size());
}

int ctlab_evalProjBack(ctlab_plan* plan, int forwardProjection, int forwardMethod, int backwardMethod, float* d_out,
                       void* stream)
{
    clear_error();
    if (forwardMethod != 0 || backwardMethod != 0) {
        set_error(CTLAB_ERR_MODE, "only the branchless models (method 0) are built");
        return CTLAB_ERR_MODE;
    }
    if (!plan || !d_out) {
        set_error(CTLAB_ERR_ARG, "need a plan and an output buffer");
        return CTLAB_ERR_ARG;
    }
    const long long nvox = (long long)plan->in.XN * plan->in.YN * plan->in.ZN;
    const long long nprj = (long long)plan->in.PN * plan->in.DNU * plan->in.DNV;
    const long long nin = forwardProjection ? nvox : nprj;
    hipStream_t s = (hipStream_t)stream;
    float* ones = nullptr;
    if (hip_check(hipMalloc(&ones, nin * sizeof(float)), "hipMalloc")) return ctlab_last_error();
    int rc = hip_check(rk_fill(ones, 1.0f, (long)nin, s), "fill") ? ctlab_last_error() : CTLAB_OK;
    if (rc == CTLAB_OK)
        rc = forwardProjection ? ctlab_proj(plan, ones, d_out, stream) : ctlab_back(plan, ones, d_out, stream);
    if (rc == CTLAB_OK && hip_check(hipStreamSynchronize(s), "sync")) rc = ctlab_last_error();
    hipFree(ones);
    return rc;
}

int ctlab_OS_SART(float* reconImg, const float* initImg, const float* hprj, const float* hangs, const float* hzPos,
                  const byte* mask, const char* volumeName, int osNum, int iterNum, int outputMedRes, float sid,
                  float sdd, int DNU, int DNV, int PN, float imgXCenter, float imgYCenter, float imgZCenter, int XN,
                  int YN, int ZN, float dx, float dz, float col_size, float row_size, float col_offset,
                  float row_offset, int forwardMethod, int backwardMethod)
{
    clear_error();
    if (!reconImg || iterNum < 0 || osNum < 1 || osNum > PN || (outputMedRes && !(volumeName && volumeName[0]))) {
        set_error(CTLAB_ERR_ARG, "need reconImg, iterNum >= 0, 1 <= osNum <= PN, and a volume name for outputMedRes");
        return CTLAB_ERR_ARG;
    }
    Driver D;
    if (driver_setup(D, initImg, hprj, hangs, hzPos, mask, osNum, sid, sdd, DNU, DNV, PN, imgXCenter, imgYCenter,
                     imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset, row_offset, forwardMethod,
                     backwardMethod))
        return ctlab_last_error();
    std::vector<float> h;
    for (int k = 1; k <= iterNum; ++k) {
        if (ctlab_os_sart(D.recon, D.d_x, 1, 1.0f, 0, 1, nullptr)) return ctlab_last_error();
        if (outputMedRes && save_volume(D, h, iteration_name(volumeName, k).c_str())) return ctlab_last_error();
    }
    if (save_volume(D, h, volumeName)) return ctlab_last_error();
    std::memcpy(reconImg, h.data(), h.size() * sizeof(float));
    return CTLAB_OK;
}

int ctlab_CG(float* reconImg, const float* initImg, const float* hprj, const float* hangs, const float* hzPos,
             const byte* mask, const char* volumeName, int osNum, int iterNum, int outputMedRes, float sid, float sdd,
             int DNU, int DNV, int PN, float imgXCenter, float imgYCenter, float imgZCenter, int XN, int YN, int ZN,
             float dx, float dz, float col_size, float row_size, float col_offset, float row_offset,
             int forwardMethod, int backwardMethod, int cgm)
{
    clear_error();
    (void)osNum;   // CG runs over all views
    if (!reconImg || iterNum < 0 || cgm < 0 || cgm > 4 || (outputMedRes && !(volumeName && volumeName[0]))) {
        set_error(CTLAB_ERR_ARG, "need reconImg, iterNum >= 0, 0 <= cgm <= 4, and a volume name for outputMedRes");
        return CTLAB_ERR_ARG;
    }
    Driver D;
    if (driver_setup(D, initImg, hprj, hangs, hzPos, mask, 1, sid, sdd, DNU, DNV, PN, imgXCenter, imgYCenter,
                     imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset, row_offset, forwardMethod,
                     backwardMethod))
        return ctlab_last_error();
    std::vector<float> h;
    const int rc = cg_run(D.recon, D.d_x, iterNum, cgm, nullptr, nullptr, [&](int k, const float*) {
        return outputMedRes ? save_volume(D, h, iteration_name(volumeName, k).c_str()) : CTLAB_OK;
    });
    if (rc) return rc;
    if (save_volume(D, h, volumeName)) return ctlab_last_error();
    std::memcpy(reconImg, h.data(), h.size() * sizeof(float));
    return CTLAB_OK;
}

}  // extern "C"

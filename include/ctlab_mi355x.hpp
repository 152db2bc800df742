/*
 * ctlab_mi355x.hpp -- C++ forms of CTLAB's host-level functions over the C ABI
 * of libctlab_mi355x.so: the reference's own names and argument lists, so
 * that C++ code written against CTLAB's headers compiles unchanged.
 *
 *   readData_host / writeData_host / writeFileFromDeviceToDisk   src/readwrite.hpp:15-113
 *   getVolZIdxPair / getPrjIdxPair                                src/utilities.cuh:478-532
 *   OS_SART / CG (+ CGMethod, ForwardDDMethod, BackwardDDMethod)  src/DD3_GPU_recon.h:14-103
 *
 * thrust::host_vector overloads are provided under hipcc (or with
 * CTLAB_WITH_THRUST defined) when <thrust/host_vector.h> is found, std::vector
 * overloads always.  Failures throw std::runtime_error
 * with the library's message (the C ABI itself never throws).
 */
#ifndef CTLAB_MI355X_HPP_
#define CTLAB_MI355X_HPP_

#include <stdexcept>
#include <string>
#include <vector>

#include "ctlab_mi355x.h"

/* rocThrust's headers compile under hipcc only */
#if (defined(__HIPCC__) || defined(CTLAB_WITH_THRUST)) && defined(__has_include)
#if __has_include(<thrust/host_vector.h>)
#include <thrust/host_vector.h>
#define CTLAB_HAVE_THRUST 1
#endif
#endif

typedef const unsigned int cuint;
enum ForwardDDMethod { PROJ_BRANCHLESS = 0, PROJ_PSEUDODISTANCE = 2 };
enum BackwardDDMethod { BACK_BRANCHLESS = 0, BACK_PSEUDODISTANCE = 2, BACK_ZLINEBRANCHLESS = 3 };
enum CGMethod { CG_FR, CG_PRP, CG_CW, CG_DI, CG_DY };

namespace ctlab_detail {
inline void check(int rc)
{
    if (rc != CTLAB_OK) throw std::runtime_error(std::string("ctlab: ") + ctlab_last_error_message());
}
template <class V> inline void sized(V& v, cuint n)
{
    if (v.size() < n) v.resize(n);
}
}  // namespace ctlab_detail

/* ---- raw file input / output (host memory, then device pointers) ------- */
inline void readData_host(const std::string& FileName, std::vector<float>& inputData, cuint FileSize)
{
    ctlab_detail::sized(inputData, FileSize);
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), inputData.data(), FileSize, 4));
}
inline void readData_host(const std::string& FileName, std::vector<double>& inputData, cuint FileSize)
{
    ctlab_detail::sized(inputData, FileSize);
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), inputData.data(), FileSize, 8));
}
inline void readData_host(const std::string& FileName, float* inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), inputData, FileSize, 4));
}
inline void readData_host(const std::string& FileName, double* inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), inputData, FileSize, 8));
}
inline void writeData_host(const std::string& FileName, const std::vector<float>& inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), inputData.data(), FileSize, 4));
}
inline void writeData_host(const std::string& FileName, const std::vector<double>& inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), inputData.data(), FileSize, 8));
}
inline void writeData_host(const std::string& FileName, float* inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), inputData, FileSize, 4));
}
inline void writeData_host(const std::string& FileName, double* inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), inputData, FileSize, 8));
}
/* device pointers; `bytes` counts elements, as the reference documents it */
inline void writeFileFromDeviceToDisk(const std::string& FileName, float* data, cuint bytes)
{
    ctlab_detail::check(ctlab_write_raw_device(FileName.c_str(), data, bytes, 4));
}
inline void writeFileFromDeviceToDisk(const std::string& FileName, double* data, cuint bytes)
{
    ctlab_detail::check(ctlab_write_raw_device(FileName.c_str(), data, bytes, 8));
}

/* ---- helical index ranges ------------------------------------------------ */
template <class Vec>
inline void getVolZIdxPair(const Vec& zPos, const size_t PrjIdx_Start, const size_t PrjIdx_End, const float detCntIdxV,
                           const float detStpZ, const int DNV, const float objCntIdxZ, const float dz, const int ZN,
                           int& ObjIdx_Start, int& ObjIdx_End)
{
    std::vector<float> z(zPos.begin(), zPos.end());
    ctlab_detail::check(ctlab_getVolZIdxPair(z.data(), (int)z.size(), (int)PrjIdx_Start, (int)PrjIdx_End, detCntIdxV,
                                             detStpZ, DNV, objCntIdxZ, dz, ZN, &ObjIdx_Start, &ObjIdx_End));
}
template <class Vec>
inline void getPrjIdxPair(const Vec& zPos, const size_t ObjZIdx_Start, const size_t ObjZIdx_End,
                          const float objCntIdxZ, const float dz, const int ZN, const float detCntIdxV,
                          const float detStpZ, const int DNV, int& prjIdx_Start, int& prjIdx_End)
{
    std::vector<float> z(zPos.begin(), zPos.end());
    ctlab_detail::check(ctlab_getPrjIdxPair(z.data(), (int)z.size(), (int)ObjZIdx_Start, (int)ObjZIdx_End, objCntIdxZ,
                                            dz, ZN, detCntIdxV, detStpZ, DNV, &prjIdx_Start, &prjIdx_End));
}

/* ---- reconstruction drivers ---------------------------------------------- */
template <class FV, class BV>
inline void OS_SART(FV& reconImg, FV& initImg, FV& hprj, FV& hangs, FV& hzPos, BV& mask, const std::string& VolumeName,
                    const int osNum, const int iterNum, const bool outputMedRes, const float sid, const float sdd,
                    const int DNU, const int DNV, const int PN, const float imgXCenter, const float imgYCenter,
                    const float imgZCenter, const int XN, const int YN, const int ZN, const float dx, const float dz,
                    const float col_size, const float row_size, const float col_offset, const float row_offset,
                    const ForwardDDMethod forwardMethod, const BackwardDDMethod backwardMethod)
{
    ctlab_detail::sized(reconImg, (cuint)((size_t)XN * YN * ZN));
    ctlab_detail::check(ctlab_OS_SART(
        &reconImg[0], initImg.empty() ? nullptr : &initImg[0], &hprj[0], &hangs[0], hzPos.empty() ? nullptr : &hzPos[0],
        mask.empty() ? nullptr : (const byte*)&mask[0], VolumeName.c_str(), osNum, iterNum, outputMedRes ? 1 : 0, sid,
        sdd, DNU, DNV, PN, imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset,
        row_offset, (int)forwardMethod, (int)backwardMethod));
}
template <class FV, class BV>
inline void CG(FV& reconImg, FV& initImg, FV& hprj, FV& hangs, FV& hzPos, BV& mask, const std::string& VolumeName,
               const int osNum, const int iterNum, const bool outputMedRes, const float sid, const float sdd,
               const int DNU, const int DNV, const int PN, const float imgXCenter, const float imgYCenter,
               const float imgZCenter, const int XN, const int YN, const int ZN, const float dx, const float dz,
               const float col_size, const float row_size, const float col_offset, const float row_offset,
               const ForwardDDMethod forwardMethod, const BackwardDDMethod backwardMethod, const CGMethod cgm)
{
    ctlab_detail::sized(reconImg, (cuint)((size_t)XN * YN * ZN));
    ctlab_detail::check(ctlab_CG(
        &reconImg[0], initImg.empty() ? nullptr : &initImg[0], &hprj[0], &hangs[0], hzPos.empty() ? nullptr : &hzPos[0],
        mask.empty() ? nullptr : (const byte*)&mask[0], VolumeName.c_str(), osNum, iterNum, outputMedRes ? 1 : 0, sid,
        sdd, DNU, DNV, PN, imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset,
        row_offset, (int)forwardMethod, (int)backwardMethod, (int)cgm));
}

#ifdef CTLAB_HAVE_THRUST
inline void readData_host(const std::string& FileName, thrust::host_vector<float>& inputData, cuint FileSize)
{
    ctlab_detail::sized(inputData, FileSize);
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), &inputData[0], FileSize, 4));
}
inline void readData_host(const std::string& FileName, thrust::host_vector<double>& inputData, cuint FileSize)
{
    ctlab_detail::sized(inputData, FileSize);
    ctlab_detail::check(ctlab_read_raw(FileName.c_str(), &inputData[0], FileSize, 8));
}
inline void writeData_host(const std::string& FileName, const thrust::host_vector<float>& inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), &inputData[0], FileSize, 4));
}
inline void writeData_host(const std::string& FileName, const thrust::host_vector<double>& inputData, cuint FileSize)
{
    ctlab_detail::check(ctlab_write_raw(FileName.c_str(), &inputData[0], FileSize, 8));
}
#endif

#endif /* CTLAB_MI355X_HPP_ */

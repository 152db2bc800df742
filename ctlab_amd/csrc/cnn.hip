// Host object of the learned proximal CNN: parameters (FP32 master + BF16
// images), activations and gradients in HBM, forward, forward+backward, Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ctlab_mi355x.h"
#include "cnn_tiles.hpp"
#include "plan.hpp"
#include "status.hpp"

namespace ctlab {
__global__ void k_conv32(const __bf16*, const __bf16*, const float*, const __bf16*, __bf16*, __bf16*, int, int,
                         int, int, int, int);
template <bool kCM>
__global__ void k_conv_first(const float*, const float*, const float*, __bf16*, __bf16*, int, int, int, int);
__global__ void k_conv_last(const __bf16*, const float*, const float*, const float*, float*, const float*, float*,
                            float, double*, int, int, int);
__global__ void k_dgrad_last(const float*, const float*, const __bf16*, __bf16*, __bf16*, int, int, int, int);
__global__ void k_wgrad32(const __bf16*, const __bf16*, float*, int, int, int, long);
__global__ void k_wgrad32_sweep(const __bf16*, const __bf16*, float*, int, int, int);
__global__ void k_wgrad_first(const __bf16*, const float*, float*, int, int, int, long);
__global__ void k_wgrad_last(const float*, const __bf16*, float*, int, int, int, long);
__global__ void k_bias_grad_partial(const __bf16*, long, long, float*, int);
__global__ void k_sum_partial(const float*, long, long, float*);
__global__ void k_reduce_chunks(const float*, int, int, int, float*);
__global__ void k_flip_transpose(const __bf16*, __bf16*);
__global__ void k_adam(float*, const float*, float*, float*, long, float, float, float, float, float, float);
__global__ void k_to_bf16(const float*, __bf16*, float*, long);
}  // namespace ctlab

using namespace ctlab;

namespace {
constexpr int CH = 32;
constexpr long kW32 = 27L * CH * CH;
// flat parameter offsets: W1 b1 W2 b2 W3 b3 W4 b4 W5 b5
constexpr long kOffW[5] = {0, 27 * CH + CH, 27 * CH + CH + kW32 + CH, 27 * CH + CH + 2 * (kW32 + CH),
                           27 * CH + CH + 3 * (kW32 + CH)};
constexpr long kLenW[5] = {27 * CH, kW32, kW32, kW32, 27 * CH};
constexpr long kLenB[5] = {CH, CH, CH, CH, 1};
constexpr long kNParams = 27 * CH + CH + 3 * (kW32 + CH) + 27 * CH + 1;
constexpr long kWgradChunk = 8192;     // voxels per wgrad partial (multiple of 128: 4 wavefronts x runs of 32)
constexpr long kVecChunk = 512;        // voxels per VALU wgrad partial
constexpr long kRedTmp = 2048 * 64;    // reduce_chunks: groups x outputs <= (2048 / ceil(n/64)) * n
constexpr long kSweepOut = kW32 + CH;  // k_wgrad32_sweep partial: a layer's W and b gradients

// k_wgrad32_sweep reads the run-blocked copies (runs of 32 z) through one
// buffer resource each (64 * nvox bytes) whose offsets, with up to two
// invalid-marks of 1.5 * 2^30 added, must stay below 2^32 and past the
// resource: ZN % 32 == 0 and 64 * nvox <= 2^30 (16.8M voxels, 512 x 512 x 64).
// Other volumes keep channel-major copies and take k_wgrad32.
bool sweep_ok(int XN, int YN, int ZN)
{
    return ZN % 32 == 0 && 64L * XN * YN * ZN <= (1L << 30);
}

// blocks of k_wgrad32_sweep: ~256 x-steps each, at most 512 (two 4-wavefront
// blocks per CU at <= 256 VGPRs), a multiple of the 8 XCDs.  Fewer, longer
// blocks also mean fewer 110 KB partials to add up (256 x 256 x 32: 256
// blocks, 137 us + 10 us of reduction; 512 blocks, 140 + 20).
long sweep_blocks(int XN, int YN, int ZN)
{
    const long steps = (long)YN * (ZN / 32) * XN;
    const long nb = std::min(512L, std::max(8L, steps / 256));
    return (nb + 7) / 8 * 8;
}
}  // namespace

struct ctlab_cnn {
    int device = 0;
    int XN = 0, YN = 0, ZN = 0, training = 0;
    long nvox = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // FP32 flat
    float* wr = nullptr;            // FP32 copy holding BF16-rounded W1 / W5
    __bf16* wb = nullptr;           // BF16 W2..W4 (3 * kW32)
    __bf16* wt = nullptr;           // flipped transposed BF16 W2..W4
    __bf16* a_cl[4] = {};           // post-ReLU activations of layers 1..4
    __bf16* a_cm[4] = {};
    __bf16* g_cl[2] = {};           // gradient ping-pong
    __bf16* g_cm[2] = {};
    float* out = nullptr;
    float* g5 = nullptr;
    float* part = nullptr;
    float* rtmp = nullptr;              // reduce_chunks group partials (kRedTmp floats)
    double* lpart = nullptr;
    long nchunks = 0, nvchunks = 0;
    long nsweep = 0;                    // k_wgrad32_sweep blocks (0: k_wgrad32 path)
    long adam_t = 0;
};

namespace {
int chk(hipError_t e, const char* what) { return hip_check(e, what); }

unsigned blocks(long n, int per) { return (unsigned)((n + per - 1) / per); }

void cnn_free(ctlab_cnn* N)
{
    if (!N) return;
    DeviceGuard dg(N->device);
    void* ptrs[] = {N->p, N->g, N->m1, N->m2, N->wr, N->wb, N->wt, N->a_cl[0], N->a_cl[1], N->a_cl[2], N->a_cl[3],
                    N->a_cm[0], N->a_cm[1], N->a_cm[2], N->a_cm[3], N->g_cl[0], N->g_cl[1], N->g_cm[0],
                    N->g_cm[1], N->out, N->g5, N->part, N->rtmp, N->lpart};
    for (void* q : ptrs)
        if (q) hipFree(q);
    delete N;
}

// refresh the kernel-side weight images from the FP32 master copy
int refresh_weights(ctlab_cnn* N, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(27 * CH, 256)), dim3(256), 0, s, N->p + kOffW[0], (__bf16*)nullptr,
                       N->wr, 27L * CH);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(27 * CH, 256)), dim3(256), 0, s, N->p + kOffW[4], (__bf16*)nullptr,
                       N->wr + 27 * CH, 27L * CH);
    for (int l = 1; l <= 3; ++l) {
        __bf16* wb = N->wb + (l - 1) * kW32;
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(kW32, 256)), dim3(256), 0, s, N->p + kOffW[l], wb, (float*)nullptr,
                           kW32);
        hipLaunchKernelGGL(k_flip_transpose, dim3(blocks(kW32, 256)), dim3(256), 0, s, wb, N->wt + (l - 1) * kW32);
    }
    return chk(hipGetLastError(), "weight refresh");
}

int forward(ctlab_cnn* N, const float* x, float* out, const float* target, hipStream_t s)
{
    const int XN = N->XN, YN = N->YN, ZN = N->ZN;
    const bool keep_cm = N->training;
    const int rb = N->nsweep > 0;       // run-blocked channel-major copies
    const int nyb = conv32_rows(XN, YN, ZN);
    // layer 1
    if (keep_cm)
        hipLaunchKernelGGL(k_conv_first<true>, dim3(conv_first_grid(XN, YN, ZN)), dim3(256), 0, s, x, N->wr,
                           N->p + kOffW[0] + kLenW[0], N->a_cl[0], N->a_cm[0], XN, YN, ZN, rb);
    else
        hipLaunchKernelGGL(k_conv_first<false>, dim3(conv_first_grid(XN, YN, ZN)), dim3(256), 0, s, x, N->wr,
                           N->p + kOffW[0] + kLenW[0], N->a_cl[0], (__bf16*)nullptr, XN, YN, ZN, rb);
    // layers 2..4 (inference ping-pongs between a_cl[0] and a_cl[1])
    for (int l = 1; l <= 3; ++l) {
        const __bf16* in = N->training ? N->a_cl[l - 1] : N->a_cl[(l - 1) & 1];
        __bf16* ocl = N->training ? N->a_cl[l] : N->a_cl[l & 1];
        __bf16* ocm = N->training ? N->a_cm[l] : nullptr;
        hipLaunchKernelGGL(k_conv32, dim3(conv32_grid(XN, YN, ZN, nyb)), dim3(512), 0, s, in, N->wb + (l - 1) * kW32,
                           N->p + kOffW[l] + kLenW[l], (const __bf16*)nullptr, ocl, ocm, XN, YN, ZN, 1, rb, nyb);
    }
    const __bf16* a4 = N->training ? N->a_cl[3] : N->a_cl[1];
    hipLaunchKernelGGL(k_conv_last, dim3(conv_last_grid(XN, YN, ZN)), dim3(256), 0, s, a4, N->wr + 27 * CH,
                       N->p + kOffW[4] + kLenW[4], x, out, target, N->g5, (float)(1.0 / (double)N->nvox),
                       target ? N->lpart : (double*)nullptr, XN, YN, ZN);
    return chk(hipGetLastError(), "cnn forward");
}

// out[i] = sum over nchunks partial rows of part (row c at part + c*n), in a
// fixed order: chunk groups spread over ~2048 blocks write partial sums to
// N->rtmp, which a second pass adds up.
void reduce_chunks(ctlab_cnn* N, const float* part, int nchunks, int n, float* out, hipStream_t s)
{
    const int gx = (n + 63) / 64;
    int G = std::max(1, std::min(2048 / gx, nchunks / 16));
    const int per = (nchunks + G - 1) / G;
    G = (nchunks + per - 1) / per;
    if (G == 1) {
        hipLaunchKernelGGL(k_reduce_chunks, dim3(gx, 1), dim3(256), 0, s, part, nchunks, nchunks, n, out);
    } else {
        hipLaunchKernelGGL(k_reduce_chunks, dim3(gx, G), dim3(256), 0, s, part, nchunks, per, n, N->rtmp);
        hipLaunchKernelGGL(k_reduce_chunks, dim3(gx, 1), dim3(256), 0, s, (const float*)N->rtmp, G, G, n, out);
    }
}

int backward(ctlab_cnn* N, const float* x, hipStream_t s)
{
    const int XN = N->XN, YN = N->YN, ZN = N->ZN;
    const long n = N->nvox;
    const int rb = N->nsweep > 0;       // run-blocked channel-major copies
    const int nyb = conv32_rows(XN, YN, ZN);
    float* G = N->g;
    // layer 5
    hipLaunchKernelGGL(k_wgrad_last, dim3((unsigned)N->nvchunks), dim3(256), 0, s, N->g5, N->a_cl[3], N->part, XN, YN,
                       ZN, kVecChunk);
    reduce_chunks(N, N->part, (int)N->nvchunks, 27 * CH, G + kOffW[4], s);
    hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)N->nvchunks), dim3(256), 0, s, N->g5, n, kVecChunk, N->part);
    reduce_chunks(N, N->part, (int)N->nvchunks, 1, G + kOffW[4] + kLenW[4], s);
    hipLaunchKernelGGL(k_dgrad_last, dim3(blocks(n, 256)), dim3(256), 0, s, N->g5, N->wr + 27 * CH, N->a_cl[3],
                       N->g_cl[0], N->g_cm[0], XN, YN, ZN, rb);
    int cur = 0;   // g_*[cur] holds dL/dz of layer l+1 (l = 3, 2, 1 below: layers 4, 3, 2)
    for (int l = 3; l >= 1; --l) {
        // weight / bias gradients of layer l+1 (index l): input a_{l} = a_cl[l-1]
        if (N->nsweep) {
            // W and b of the layer in one pass (b follows W in the flat vector)
            hipLaunchKernelGGL(k_wgrad32_sweep, dim3((unsigned)N->nsweep), dim3(256), 0, s, N->g_cm[cur],
                               N->a_cm[l - 1], N->part, XN, YN, ZN);
            reduce_chunks(N, N->part, (int)N->nsweep, (int)kSweepOut, G + kOffW[l], s);
        } else {
            hipLaunchKernelGGL(k_wgrad32, dim3((unsigned)N->nchunks, 9), dim3(256), 0, s, N->g_cm[cur],
                               N->a_cm[l - 1], N->part, XN, YN, ZN, kWgradChunk);
            reduce_chunks(N, N->part, (int)N->nchunks, (int)kW32, G + kOffW[l], s);
            hipLaunchKernelGGL(k_bias_grad_partial, dim3((unsigned)N->nvchunks), dim3(256), 0, s, N->g_cm[cur], n,
                               kVecChunk, N->part, 0);
            reduce_chunks(N, N->part, (int)N->nvchunks, CH, G + kOffW[l] + kLenW[l], s);
        }
        // input gradient through layer l+1, masked by the ReLU of layer l
        hipLaunchKernelGGL(k_conv32, dim3(conv32_grid(XN, YN, ZN, nyb)), dim3(512), 0, s, N->g_cl[cur],
                           N->wt + (l - 1) * kW32, (const float*)nullptr, N->a_cl[l - 1], N->g_cl[cur ^ 1],
                           N->g_cm[cur ^ 1], XN, YN, ZN, 0, rb, nyb);
        cur ^= 1;
    }
    // layer 1
    hipLaunchKernelGGL(k_wgrad_first, dim3((unsigned)N->nvchunks), dim3(256), 0, s, N->g_cl[cur], x, N->part, XN, YN,
                       ZN, kVecChunk);
    reduce_chunks(N, N->part, (int)N->nvchunks, 27 * CH, G + kOffW[0], s);
    hipLaunchKernelGGL(k_bias_grad_partial, dim3((unsigned)N->nvchunks), dim3(256), 0, s, N->g_cm[cur], n, kVecChunk,
                       N->part, rb);
    reduce_chunks(N, N->part, (int)N->nvchunks, CH, G + kOffW[0] + kLenW[0], s);
    return chk(hipGetLastError(), "cnn backward");
}
}  // namespace

extern "C" {

int ctlab_cnn_num_params(void) { return (int)kNParams; }

ctlab_cnn* ctlab_cnn_create(int XN, int YN, int ZN, int training, int gpunum)
{
    clear_error();
    if (XN <= 0 || YN <= 0 || ZN <= 0) {
        set_error(CTLAB_ERR_ARG, "sizes must be positive");
        return nullptr;
    }
    if (training && ZN % 8 != 0) {
        set_error(CTLAB_ERR_ARG, "training needs ZN %% 8 == 0 (got %d)", ZN);
        return nullptr;
    }
    // one-dimensional launch grids of the tiled convolution kernels; the 32 -> 32
    // layers address one y row of channels-last activations (64 bytes per voxel)
    // with 32-bit offsets
    if (conv32_tiles(XN, YN, ZN, 2) > 0x7FFFFFF0L || conv_last_tiles(XN, YN, ZN) > 0x7FFFFFF0L ||
        conv_first_tiles(XN, YN, ZN) > 0x7FFFFFF0L || (long)XN * ZN * 64 > 0x7FFFFFFFL) {
        set_error(CTLAB_ERR_ARG, "volume %d x %d x %d exceeds the convolution kernels' grid", XN, YN, ZN);
        return nullptr;
    }
    ctlab_cnn* N = new ctlab_cnn();
    N->device = gpunum;
    N->XN = XN; N->YN = YN; N->ZN = ZN; N->training = training;
    N->nvox = (long)XN * YN * ZN;
    if (chk(hipSetDevice(gpunum), "hipSetDevice")) { delete N; return nullptr; }
    DeviceGuard dg(gpunum);
    const size_t P = kNParams * sizeof(float);
    const size_t A = (size_t)N->nvox * CH * sizeof(__bf16);
    N->nchunks = (N->nvox + kWgradChunk - 1) / kWgradChunk;
    N->nvchunks = (N->nvox + kVecChunk - 1) / kVecChunk;
    N->nsweep = training && sweep_ok(XN, YN, ZN) ? sweep_blocks(XN, YN, ZN) : 0;
    bool ok = !chk(hipMalloc(&N->p, P), "hipMalloc") && !chk(hipMalloc(&N->wr, 2 * 27 * CH * sizeof(float)), "hipMalloc") &&
              !chk(hipMalloc(&N->wb, 3 * kW32 * sizeof(__bf16)), "hipMalloc") &&
              !chk(hipMalloc(&N->wt, 3 * kW32 * sizeof(__bf16)), "hipMalloc") &&
              !chk(hipMalloc(&N->out, N->nvox * sizeof(float)), "hipMalloc") &&
              !chk(hipMalloc(&N->g5, N->nvox * sizeof(float)), "hipMalloc") &&
              !chk(hipMalloc(&N->lpart, conv_last_tiles(XN, YN, ZN) * sizeof(double)), "hipMalloc");
    const int nact = training ? 4 : 2;
    for (int i = 0; ok && i < nact; ++i) ok = !chk(hipMalloc(&N->a_cl[i], A), "hipMalloc(activations)");
    if (training) {
        for (int i = 0; ok && i < 4; ++i) ok = !chk(hipMalloc(&N->a_cm[i], A), "hipMalloc(activations)");
        for (int i = 0; ok && i < 2; ++i)
            ok = !chk(hipMalloc(&N->g_cl[i], A), "hipMalloc(gradients)") && !chk(hipMalloc(&N->g_cm[i], A), "hipMalloc(gradients)");
        const size_t part = std::max({(size_t)N->nchunks * kW32, (size_t)N->nvchunks * 27 * CH,
                                      (size_t)N->nsweep * kSweepOut}) * sizeof(float);
        ok = ok && !chk(hipMalloc(&N->g, P), "hipMalloc") && !chk(hipMalloc(&N->m1, P), "hipMalloc") &&
             !chk(hipMalloc(&N->m2, P), "hipMalloc") && !chk(hipMalloc(&N->part, part), "hipMalloc") &&
             !chk(hipMalloc(&N->rtmp, kRedTmp * sizeof(float)), "hipMalloc") &&
             !chk(hipMemset(N->m1, 0, P), "memset") && !chk(hipMemset(N->m2, 0, P), "memset");
    }
    if (!ok) { cnn_free(N); return nullptr; }
    return N;
}

void ctlab_cnn_destroy(ctlab_cnn* N) { cnn_free(N); }

int ctlab_cnn_set_params(ctlab_cnn* N, const float* h_params)
{
    clear_error();
    if (!N || !h_params) { set_error(CTLAB_ERR_ARG, "NULL cnn or params"); return CTLAB_ERR_ARG; }
    DeviceGuard dg(N->device);
    if (chk(hipMemcpy(N->p, h_params, kNParams * sizeof(float), hipMemcpyHostToDevice), "H2D")) return ctlab_last_error();
    if (refresh_weights(N, nullptr) || chk(hipDeviceSynchronize(), "refresh")) return ctlab_last_error();
    N->adam_t = 0;
    if (N->training && (chk(hipMemset(N->m1, 0, kNParams * sizeof(float)), "memset") ||
                        chk(hipMemset(N->m2, 0, kNParams * sizeof(float)), "memset")))
        return ctlab_last_error();
    return CTLAB_OK;
}

int ctlab_cnn_get_params(ctlab_cnn* N, float* h_params)
{
    clear_error();
    if (!N || !h_params) { set_error(CTLAB_ERR_ARG, "NULL cnn or params"); return CTLAB_ERR_ARG; }
    DeviceGuard dg(N->device);
    return chk(hipMemcpy(h_params, N->p, kNParams * sizeof(float), hipMemcpyDeviceToHost), "D2H") ? ctlab_last_error() : 0;
}

int ctlab_cnn_forward(ctlab_cnn* N, const float* d_x, float* d_out, void* stream)
{
    clear_error();
    if (!N || !d_x || !d_out) { set_error(CTLAB_ERR_ARG, "NULL cnn or buffer"); return CTLAB_ERR_ARG; }
    DeviceGuard dg(N->device);
    return forward(N, d_x, d_out, nullptr, (hipStream_t)stream) ? ctlab_last_error() : CTLAB_OK;
}

// loss = 0.5 * mean((x + net(x) - target)^2); gradients stay on the device
// (copied to h_grad when given); d_out (optional) receives x + net(x)
double ctlab_cnn_loss_grad(ctlab_cnn* N, const float* d_x, const float* d_target, float* d_out, float* h_grad,
                           void* stream)
{
    clear_error();
    if (!N || !N->training || !d_x || !d_target) {
        set_error(CTLAB_ERR_ARG, "need a training cnn, x and target");
        return NAN;
    }
    DeviceGuard dg(N->device);
    hipStream_t s = (hipStream_t)stream;
    if (forward(N, d_x, d_out ? d_out : N->out, d_target, s) || backward(N, d_x, s)) return NAN;
    const long nb = conv_last_tiles(N->XN, N->YN, N->ZN);       // one partial per k_conv_last tile
    std::vector<double> lp(nb);
    if (chk(hipMemcpyAsync(lp.data(), N->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, s), "D2H")) return NAN;
    if (h_grad && chk(hipMemcpyAsync(h_grad, N->g, kNParams * sizeof(float), hipMemcpyDeviceToHost, s), "D2H"))
        return NAN;
    if (chk(hipStreamSynchronize(s), "sync")) return NAN;
    double r2 = 0.0;
    for (double v : lp) r2 += v;
    return 0.5 * r2 / (double)N->nvox;
}

int ctlab_cnn_adam_step(ctlab_cnn* N, float lr, float beta1, float beta2, float eps, void* stream)
{
    clear_error();
    if (!N || !N->training) { set_error(CTLAB_ERR_ARG, "need a training cnn"); return CTLAB_ERR_ARG; }
    DeviceGuard dg(N->device);
    hipStream_t s = (hipStream_t)stream;
    N->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)N->adam_t));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)N->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(blocks(kNParams, 256)), dim3(256), 0, s, N->p, N->g, N->m1, N->m2, kNParams, lr,
                       beta1, beta2, eps, c1, c2);
    if (chk(hipGetLastError(), "adam") || refresh_weights(N, s)) return ctlab_last_error();
    return CTLAB_OK;
}

}  // extern "C"

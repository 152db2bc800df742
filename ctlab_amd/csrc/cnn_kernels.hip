// Learned proximal CNN kernels (gfx950): 5 layers of 3x3x3 convolution,
// 1 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after layers 1-4, residual output
// x + net(x).  BF16 weights and activations, FP32 accumulation.  No
// counterpart in CTLAB.
//
// Layouts (voxel v = (y*XN + x)*ZN + z, the volume's (Z,X,Y) order):
//   channels-last  act_cl[v*32 + c]   (MFMA A-operand: 8 channels of a voxel = 16 B)
//   channel-major  act_cm[c*nvox + v] (wgrad operands: 8 voxels of a channel = 16 B)
//   run-blocked    act_rb[(v/32)*1024 + c*32 + v%32]  (ZN % 32 == 0: the same
//                  operands, with the 32 channels of a run of 32 voxels in one 2 KB
//                  block, so a wavefront's operand load is one contiguous KB)
//   weights        w[t][co][ci], tap t = (dy+1)*9 + (dx+1)*3 + (dz+1)
// Hidden 32 -> 32 layers are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
// one MFMA per (16 voxels, 16 output channels, tap) with K = the 32 input
// channels.  The input gradient runs the same kernel with flipped, transposed
// weights and the ReLU mask fused into its epilogue; the weight gradient runs
// MFMAs over K = voxels on the channel-major copies.  The 32 -> 1 output
// layer runs on MFMA with the (dy, dx) taps as the N dimension, and so does
// the 1 -> 32 first layer (its FP32 input split into three exact BF16 parts);
// the thin gradient kernels run on the VALU.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cnn_tiles.hpp"

namespace ctlab {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int CH = 32;

__device__ __forceinline__ int tap_off(int t, int XN, int ZN)
{
    const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
    return dz + dx * ZN + dy * XN * ZN;
}

__device__ __forceinline__ bool tap_ok(int t, int y, int x, int z, int YN, int XN, int ZN)
{
    const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
    return (unsigned)(y + dy) < (unsigned)YN && (unsigned)(x + dx) < (unsigned)XN &&
           (unsigned)(z + dz) < (unsigned)ZN;
}

__device__ __forceinline__ float bf(const __bf16 v) { return (float)v; }

// element (channel c, voxel v) of a channel-major (rb = 0) or run-blocked (rb = 1) copy
__device__ __forceinline__ long cm_index(int c, long v, long nvox, int rb)
{
    return rb ? ((v >> 5) << 10) + ((long)c << 5) + (v & 31) : (long)c * nvox + v;
}

// ---------------------------------------------------------------------------
// 32 -> 32 convolution, MFMA (implicit GEMM: voxels x 27 taps x 32 channels).
//   out = act(conv(in; w) + bias) ; act: relu (relu=1) or identity
//   mask (optional, channels-last): out *= (mask > 0)   [input-gradient pass]
// Writes channels-last and (optionally) channel-major copies.
// Block = 8 wavefronts, one per CU (LDS); output tile = 8 x-columns x 32 z,
// for nyb consecutive y rows, taken two rows at a time: wavefronts 0-3 compute
// row y, 4-7 row y + 1 (each wavefront two x-columns, 4 blocks of 16 voxels
// along z, 8 accumulator chains).  All 27 taps' weights stay in LDS for the
// block's life.  The input planes (bricks: 10 columns x 34 z of one y row, all
// 32 channels, zero outside the volume) pass through a ring of three LDS
// slots.  A pair of rows takes three steps; at step k the row-y group
// multiplies plane y - 1 + k and the row-(y+1) group plane y + k, both with
// weight plane k, so the two groups share every brick and a pair needs two new
// bricks: one brick per output row, where one row per block with per-plane
// weight staging took three bricks and three weight planes.  Bricks arrive by
// LDS DMA (buffer_load_dwordx4 ... lds), issued one step before they are read,
// with one barrier per step.  The DMA writes lane i of an instruction at
// M0 + 16 i, so an instruction fills 64 consecutive rows of one chunk plane;
// loads outside the row's buffer range write zeros, which gives the halo.
// The ring slots are separate LDS arrays read without a select, so that the
// compiler can tell a slot being filled from the ones being read and issues
// no vmcnt wait between the DMA and the MFMA operand reads.
// ---------------------------------------------------------------------------
constexpr int kBX = kTX + 2, kBZ = kTZ + 2;         // input brick of one plane
constexpr int kBrick = kBX * kBZ;                   // 340 voxels

// LDS tiles are stored as chunk planes: 16-byte chunk q (channels 8q..8q+7)
// of row p (a voxel of the brick, or an output channel of a tap's weights) at
// slot q * stride + p, so lane (m, kq) of an MFMA operand reads chunk kq of
// row base + m and a tap's shift stays a constant offset.  Measured on gfx950
// (ds_read_b128, 64 lanes as here): a read takes 4 LDS cycles when the plane
// stride is a multiple of 256 bytes, and 8 (2-way bank conflicts) with 64-byte
// rows (the layout [row][chunk]) or with a plane stride of 5440, 4160 or 4224
// bytes.  Register-staged planes are filled in the same (row, chunk) lane
// pattern they are read in: staging index i holds row 16 * (i / 64) + i % 16,
// chunk (i / 16) % 4.
constexpr int kBrickP = 352;                        // brick plane stride: 5,632 B (register-staged)
constexpr int kDmaP = 384;                          // brick plane stride for DMA: 6 x 64 rows (6,144 B)
constexpr int kSlot = 4 * kDmaP;                    // one DMA brick: 24,576 B
constexpr int kWPA = 27 * CH;                       // weight plane stride, all taps: 864 rows (13,824 B)

__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
k_conv32(const __bf16* __restrict__ in, const __bf16* __restrict__ w, const float* __restrict__ bias,
         const __bf16* __restrict__ mask, __bf16* __restrict__ out_cl, __bf16* __restrict__ out_cm,
         int XN, int YN, int ZN, int relu, int rb, int nyb)
{
    __shared__ bf16x8 sW[4 * kWPA];                 // 55,296 B: all 27 taps
    __shared__ bf16x8 sB0[kSlot], sB1[kSlot], sB2[kSlot];   // 73,728 B: ring of three bricks
    __shared__ bf16x8 sO8[8 * 64 * CH / 8];         // 32,768 B: each wavefront's output patch
    const long ntiles = conv32_tiles(XN, YN, ZN, nyb);
    // contiguous tile ranges per XCD (blocks are dealt to XCDs round-robin); the grid
    // is 8 * ceil(ntiles / 8), so (xcd, j) covers [0, 8 * per) once.  Tile order: z,
    // then x, then y range, so an XCD's concurrent blocks share their halo columns in L2.
    const long per = (long)gridDim.x >> 3;
    const long tile = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;                     // whole block: before any barrier
    const int ntz = (ZN + kTZ - 1) / kTZ, ntx = (XN + kTX - 1) / kTX;
    const int zt = (int)(tile % ntz);
    const long r = tile / ntz;
    const int x0 = (int)(r % ntx) * kTX, z0 = zt * kTZ;
    const int y0 = (int)(r / ntx) * nyb, ny = min(nyb, YN - y0);

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = wv >> 2, w4 = wv & 3;           // row y + grp; x-columns 2 w4, 2 w4 + 1
    const int m = lane & 15, kq = lane >> 4;
    // block b of this wavefront: column 2*w4 + (b >> 1), z 16*(b & 1) + m; its
    // centre voxel in the brick is (column + 1, z + 1)
    int pb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) pb[b] = (2 * w4 + (b >> 1) + 1) * kBZ + 16 * (b & 1) + m + 1;

    // A brick is 24 DMA instructions (chunk plane q, rows 64 j .. 64 j + 63); wavefront
    // wv issues numbers wv, wv + 8, wv + 16.  Lane offsets are taken within one y row of
    // the input, whose buffer range is the row; rows past the brick, outside the volume
    // in x or z, or of a y row outside it read past the range and write zeros.
    const unsigned rowbytes = (unsigned)XN * (unsigned)ZN * 64u;
    unsigned loff[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int I = wv + 8 * e, q = I / 6, p = 64 * (I % 6) + lane;
        const int gx = x0 - 1 + p / kBZ, gz = z0 - 1 + p % kBZ;
        const bool ok = p < kBrick && (unsigned)gx < (unsigned)XN && (unsigned)gz < (unsigned)ZN;
        loff[e] = ok ? (unsigned)(((gx * ZN + gz) * 4 + q) * 16) : rowbytes;
    }
    auto load_brick = [&](int n, bf16x8* slot) {    // plane n: input row y0 - 1 + n
        const int yy = y0 - 1 + n;
        const bool row_in = (unsigned)yy < (unsigned)YN;
        const __bf16* base = in + (row_in ? (long)yy * XN * ZN * CH : 0L);
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, row_in ? rowbytes : 0u, 0x00020000);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int I = wv + 8 * e;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rs, (__attribute__((address_space(3))) void*)(slot + (I / 6) * kDmaP + 64 * (I % 6)), 16, loff[e], 0, 0, 0);
        }
    };
    load_brick(0, sB0);
    load_brick(1, sB1);
    auto srow = [](int i) { return ((i >> 6) << 4) | (i & 15); };
    auto schunk = [](int i) { return (i >> 4) & 3; };
    for (int i = threadIdx.x; i < 27 * CH * CH / 8; i += 512)
        sW[schunk(i) * kWPA + srow(i)] = ((const bf16x8*)w)[srow(i) * 4 + schunk(i)];

    const long nvox = (long)XN * YN * ZN;
    __bf16* sO = (__bf16*)sO8 + wv * 64 * CH;       // 64 voxels x 32 channels (4 KB) of this wavefront
    const float bias0 = bias ? bias[m] : 0.0f, bias1 = bias ? bias[16 + m] : 0.0f;
    f32x4 acc[4][2];
    auto taps = [&](const bf16x8* sA, int k) {      // the 9 taps of weight plane k (dy = k - 1) on brick sA
#pragma unroll
        for (int tt = 0; tt < 9; ++tt) {
            const int shift = (tt / 3 - 1) * kBZ + (tt % 3 - 1);      // (dx, dz) in the brick
            const int wr = (9 * k + tt) * CH;
            const bf16x8 b0 = sW[kq * kWPA + wr + m];
            const bf16x8 b1 = sW[kq * kWPA + wr + 16 + m];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bf16x8 a = sA[kq * kDmaP + pb[b] + shift];
                acc[b][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[b][0], 0, 0, 0);
                acc[b][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[b][1], 0, 0, 0);
            }
        }
    };
    auto step_sync = [&]() {
        __builtin_amdgcn_s_waitcnt(0x0f70);         // vmcnt(0): this wavefront's DMA has landed
        __syncthreads();                            // everyone's; the slot to be refilled is free
    };
    // Epilogue of row y.  The D fragments (lane: 4 voxels of one output channel) go,
    // with bias and ReLU, into this wavefront's [voxel][channel] patch; each lane then
    // writes 16-byte chunks (8 channels of one voxel) of the channels-last output,
    // reading the ReLU mask of the input-gradient pass in the same chunks.  The patch
    // is the wavefront's own and LDS executes one wavefront's accesses in issue order,
    // so it needs no barrier.  Writing the fragments straight out took 64 two-byte
    // stores per lane, ~20 % of the kernel at 512x512x64.
    auto epilogue = [&](int y) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int co = nb * 16 + m;
                const float bco = nb ? bias1 : bias0;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float rr = acc[b][nb][i] + bco;
                    if (relu) rr = fmaxf(rr, 0.0f);
                    sO[(16 * b + 4 * kq + i) * CH + co] = (__bf16)rr;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = lane + 64 * k;            // chunk: voxel c / 4, channels 8 * (c % 4) ..
            const int vb = c >> 2, q = c & 3;
            const int x = x0 + 2 * w4 + (vb >> 5);
            const int z = z0 + (vb & 31);
            if (x >= XN || z >= ZN) continue;
            const long v = ((long)y * XN + x) * ZN + z;
            bf16x8 o = *(const bf16x8*)(sO + vb * CH + 8 * q);
            if (mask) {
                const bf16x8 mk = *(const bf16x8*)(mask + v * CH + 8 * q);
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (!(bf(mk[e]) > 0.0f)) o[e] = (__bf16)0.0f;
                if (out_cm) *(bf16x8*)(sO + vb * CH + 8 * q) = o;   // the masked values for the copy below
            }
            *(bf16x8*)(out_cl + v * CH + 8 * q) = o;
        }
        if (out_cm) {
            // channel-major (or run-blocked) copy: lane = (channel m of block nb, 4 voxels)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int x = x0 + 2 * w4 + (b >> 1);
                const int zb = z0 + 16 * (b & 1) + 4 * kq;
                if (x >= XN) continue;
                const long v0 = ((long)y * XN + x) * ZN + zb;
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const int co = nb * 16 + m;
                    __bf16 o[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = sO[(16 * b + 4 * kq + i) * CH + co];
                    if (zb + 3 < ZN && (ZN & 3) == 0) {     // v0 is a multiple of 4: 8-byte aligned vector store
                        *(bf16x4*)(out_cm + cm_index(co, v0, nvox, rb)) = (bf16x4){o[0], o[1], o[2], o[3]};
                    } else {
                        for (int i = 0; i < 4; ++i)
                            if (zb + i < ZN) out_cm[cm_index(co, v0 + i, nvox, rb)] = o[i];
                    }
                }
            }
        }
    };
    // Pair t (rows y0 + 2t, y0 + 2t + 1): planes 2t in P and 2t + 1 in Q; plane 2t + 2 is
    // loaded into R at step 0 (R last held plane 2t - 1, read at the previous pair's
    // step 1) and plane 2t + 3 into P at step 1 (P's plane was last read at step 0).
    // The loads past the last plane read rows beyond the tile, harmlessly.  Each group
    // reads its slot through a wave-uniform branch rather than a selected pointer.
    auto pair = [&](int t, bf16x8* P, bf16x8* Q, bf16x8* R) {
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b][0] = acc[b][1] = (f32x4){0.f, 0.f, 0.f, 0.f};
        step_sync();
        load_brick(2 * t + 2, R);
        if (grp) taps(Q, 0); else taps(P, 0);
        step_sync();
        load_brick(2 * t + 3, P);
        if (grp) taps(R, 1); else taps(Q, 1);
        step_sync();
        if (grp) taps(P, 2); else taps(R, 2);
        const int y = y0 + 2 * t + grp;
        if (y < y0 + ny) epilogue(y);
    };
    const int npair = (ny + 1) >> 1;
    for (int t = 0; t < npair; t += 3) {            // slots rotate with period three pairs
        pair(t, sB0, sB1, sB2);
        if (t + 1 < npair) pair(t + 1, sB2, sB0, sB1);
        if (t + 2 < npair) pair(t + 2, sB1, sB2, sB0);
    }
}

// A wavefront's output patch in LDS: 64 voxels x 32 channels, 64 bytes per
// voxel, with the 16-byte chunk q of voxel vb at chunk q ^ ((vb >> 2) & 3), so
// that the 8-byte fragment writes (lanes: 16 voxels x 2 halves of a chunk)
// fall on distinct banks while each voxel's 64 bytes stay contiguous.
__device__ __forceinline__ int patch_off(int vb, int c)
{
    return vb * CH + ((((c >> 3) ^ (vb >> 2)) & 3) << 3) + (c & 7);
}

// ---------------------------------------------------------------------------
// first layer 1 -> 32 on MFMA: out = relu(conv(x; w1) + b1); w1 FP32 [27][32]
// holding BF16-rounded values.  Per v_mfma_f32_16x16x32_bf16: M = 16 voxels,
// N = 16 output channels, K = the 27 taps (padded to 32).  The FP32 input
// enters as three BF16 parts, x = hi + mid + lo with hi = bf16(x), mid =
// bf16(x - hi), lo = x - hi - mid; each part holds 8 significant bits, so the
// split is exact, every product with a BF16 weight is exact, and the FP32
// accumulation gives the FP32 sum of the former VALU kernel up to the order of
// the additions (that kernel took ~1,100 VALU instructions per voxel, 0.62 ms
// at 512x512x64).  A block computes the k_conv32 tile (8 x-columns x 32 z;
// wavefront w: columns 2w, 2w+1, 4 blocks of 16 voxels) for kFY consecutive y
// rows: the input planes (10 columns x 34 z of one y row, zero outside the
// volume) sit in a ring of three LDS slots, and plane y + 2 is read into
// registers while row y is computed, so each input plane is fetched once per
// block and the weight fragments once per kFY rows (one row per block with all
// three planes staged before the MFMAs: 0.57 ms at 512x512x64, the load latency
// exposed).  The output goes through LDS in 16-byte chunks as in k_conv32.
// ---------------------------------------------------------------------------
// kCM: also write the channel-major copy out_cm (training); a separate instance,
// since its address arithmetic, held across the row loop, costs 41 registers.
template <bool kCM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_conv_first(const float* __restrict__ x, const float* __restrict__ w1, const float* __restrict__ b1,
             __bf16* __restrict__ out_cl, __bf16* __restrict__ out_cm, int XN, int YN, int ZN, int rb)
{
    __shared__ float sX[3 * kBrick];                        // 4,080 B: planes y-1, y, y+1 (ring)
    __shared__ bf16x8 sO8[4 * 64 * CH / 8];                 // 16 KB: each wavefront's 64 voxels x 32 channels
    const long ntiles = conv_first_tiles(XN, YN, ZN);
    const long per = (long)gridDim.x >> 3;
    const long tile = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;                             // whole block: before any barrier
    const int ntz = (ZN + kTZ - 1) / kTZ, ntx = (XN + kTX - 1) / kTX;
    const int zt = (int)(tile % ntz);
    const long r = tile / ntz;
    const int x0 = (int)(r % ntx) * kTX, z0 = zt * kTZ;
    const int y0 = (int)(r / ntx) * kFY, ny = min(kFY, YN - y0);

    // plane element p (column p / kBZ, z p % kBZ of the brick) of row yy, 0 outside the volume
    auto element = [&](int yy, int p) {
        const int gx = x0 - 1 + p / kBZ, gz = z0 - 1 + p % kBZ;
        float v = 0.0f;
        if (p < kBrick && (unsigned)yy < (unsigned)YN && (unsigned)gx < (unsigned)XN && (unsigned)gz < (unsigned)ZN)
            v = x[((long)yy * XN + gx) * ZN + gz];
        return v;
    };
    constexpr int kNP = (kBrick + 255) / 256;               // plane elements per thread
    // slot of plane yy: (yy - y0 + 1) % 3; planes y0 - 1 and y0 go straight in
#pragma unroll
    for (int j = 0; j < kNP; ++j) {
        const int p = threadIdx.x + 256 * j;
        if (p < kBrick) {
            sX[p] = element(y0 - 1, p);
            sX[kBrick + p] = element(y0, p);
        }
    }
    // planes y0 + 1 and y0 + 2 in registers; row y then writes plane y + 1 to its slot and
    // fetches plane y + 3, which has two rows' time to arrive
    float pfA[kNP], pfB[kNP];
#pragma unroll
    for (int j = 0; j < kNP; ++j) {
        pfA[j] = element(y0 + 1, threadIdx.x + 256 * j);
        pfB[j] = element(y0 + 2, threadIdx.x + 256 * j);
    }

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    // B fragments (lane: taps 8kq .. 8kq+7 of output channel 16nb + m; taps 27..31 are 0)
    bf16x8 B[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 8 * kq + e;
            B[nb][e] = (__bf16)(t < 27 ? w1[t * CH + nb * 16 + m] : 0.0f);
        }
    // this lane's 4 output channels 16nb + 4kq + i (weights as the M operand: D[channel][voxel])
    float bc[2][4];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) bc[nb][i] = b1[16 * nb + 4 * kq + i];
    const long nvox = (long)XN * YN * ZN;
    __bf16* sO = (__bf16*)sO8 + wv * 64 * CH;

    auto row = [&](int s, float (&pf)[kNP]) {
        const int y = y0 + s;
        __syncthreads();                                    // row y - 1's gathers are done with its slot
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            const int p = threadIdx.x + 256 * j;
            if (p < kBrick) sX[((s + 2) % 3) * kBrick + p] = pf[j];
        }
        __syncthreads();
        if (s + 2 < ny) {                                   // plane y + 3, in flight during rows y, y + 1
#pragma unroll
            for (int j = 0; j < kNP; ++j) pf[j] = element(y + 3, threadIdx.x + 256 * j);
        }
        // lane's 8 taps: offset of tap 8kq + e from a voxel's centre (the centre in slot (s + 1) % 3)
        int off[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = min(8 * kq + e, 26);
            const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
            off[e] = ((s + 1 + dy + 3) % 3) * kBrick + dx * kBZ + dz;
        }

        // block b: column 2wv + (b >> 1), z 16(b & 1) + m.  Its fragments (lane: channels
        // 16nb + 4kq + i of voxel m) go with bias and ReLU into this wavefront's
        // [voxel][channel] patch, then 16-byte chunks out.  LDS executes one wavefront's
        // accesses in issue order, so the patch needs no barrier between rows.
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int ctr = (2 * wv + (b >> 1) + 1) * kBZ + 16 * (b & 1) + m + 1;   // centre in the brick
            bf16x8 hi, mid, lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float a = 8 * kq + e < 27 ? sX[ctr + off[e]] : 0.0f;
                hi[e] = (__bf16)a;
                const float r1 = a - (float)hi[e];
                mid[e] = (__bf16)r1;
                lo[e] = (__bf16)(r1 - (float)mid[e]);
            }
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[nb], hi, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[nb], mid, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[nb], lo, acc, 0, 0, 0);
                bf16x4 o4;
#pragma unroll
                for (int i = 0; i < 4; ++i) o4[i] = (__bf16)fmaxf(acc[i] + bc[nb][i], 0.0f);
                *(bf16x4*)(sO + patch_off(16 * b + m, 16 * nb + 4 * kq)) = o4;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = lane + 64 * k;                    // chunk: voxel c / 4, channels 8 * (c % 4) ..
            const int vb = c >> 2, q = c & 3;
            const int xx = x0 + 2 * wv + (vb >> 5);
            const int z = z0 + (vb & 31);
            if (xx >= XN || z >= ZN) continue;
            const long v = ((long)y * XN + xx) * ZN + z;
            *(bf16x8*)(out_cl + v * CH + 8 * q) = *(const bf16x8*)(sO + patch_off(vb, 8 * q));
        }
        if (kCM) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int xx = x0 + 2 * wv + (b >> 1);
                const int zb = z0 + 16 * (b & 1) + 4 * kq;
                if (xx >= XN) continue;
                const long v0 = ((long)y * XN + xx) * ZN + zb;
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const int co = nb * 16 + m;
                    __bf16 o[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = sO[patch_off(16 * b + 4 * kq + i, co)];
                    if (zb + 3 < ZN && (ZN & 3) == 0) {
                        *(bf16x4*)(out_cm + cm_index(co, v0, nvox, rb)) = (bf16x4){o[0], o[1], o[2], o[3]};
                    } else {
                        for (int i = 0; i < 4; ++i)
                            if (zb + i < ZN) out_cm[cm_index(co, v0 + i, nvox, rb)] = o[i];
                    }
                }
            }
        }
    };
    for (int s = 0; s < ny; s += 2) {                       // two rows per trip: pfA, pfB stay registers
        row(s, pfA);
        if (s + 1 < ny) row(s + 1, pfB);
    }
}
template __global__ void k_conv_first<false>(const float*, const float*, const float*, __bf16*, __bf16*, int, int, int, int);
template __global__ void k_conv_first<true>(const float*, const float*, const float*, __bf16*, __bf16*, int, int, int, int);

// ---------------------------------------------------------------------------
// last layer 32 -> 1, MFMA: out = x + conv(a4; w5) + b5; training forward
// also writes g5 = (out - target) * gscale and the block's sum of squares.
// Block = 4 wavefronts, output tile = 8 x-columns x 32 z, for kLY consecutive
// y rows.  Each input column of a row (the tile's 10 columns with the halo) and
// 16-voxel z block gives, in one MFMA chain over dz (K = 32 channels each),
// its shares of the 9 output columns it reaches: B column n = (dy, dx) holds
// the weights of tap (dy, dx, dz), so D[z][n] is input column c's share of
// output column c - (dy, dx), in output row (input row) - dy.  Every share has
// its own LDS slot (row, n, output column, z), in a ring of three output rows;
// after input row y + 1 row y is complete and each of its outputs adds its 9
// shares in a fixed order (deterministic).  The block walks the input rows
// y0 - 1 .. y0 + kLY once: a row (10 columns x 34 z x 32 channels, zero
// outside the volume) is staged in LDS as k_conv32's brick, the next one read
// into registers meanwhile, so each input voxel leaves HBM about once (tiles of
// 8 x 4 output columns, each reading its own 10 x 6 input columns: 2.1 GB at
// 512x512x64, 0.53 ms).  w5 holds the BF16-rounded weights as FP32, so the B
// operand is exact.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_conv_last(const __bf16* __restrict__ in, const float* __restrict__ w5, const float* __restrict__ b5,
            const float* __restrict__ x, float* __restrict__ out, const float* __restrict__ target,
            float* __restrict__ g5, float gscale, double* __restrict__ part, int XN, int YN, int ZN)
{
    constexpr int kIX = kLX + 2;                            // 10 input columns
    constexpr int kRow = 9 * kLX * kTZ;                     // one output row's shares (n, column, z)
    __shared__ bf16x8 sA[4 * kBrickP];                      // 22,528 B: the input row
    __shared__ f32x4 sD[3 * kRow / 4];                      // 27,648 B: three output rows' shares
    __shared__ double sred[4];
    const long ntiles = conv_last_tiles(XN, YN, ZN);
    const long per = (long)gridDim.x >> 3;                  // contiguous tile ranges per XCD
    const long tile = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;                             // whole block: before any barrier
    const int ntz = (ZN + kTZ - 1) / kTZ, ntx = (XN + kLX - 1) / kLX;
    const int zt = (int)(tile % ntz);
    const long r = tile / ntz;
    const int x0 = (int)(r % ntx) * kLX, y0 = (int)(r / ntx) * kLY, z0 = zt * kTZ;
    const int ny = min(kLY, YN - y0);

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    // B fragments (lane: channels 8kq..8kq+7 of column n = m), one per dz
    bf16x8 B[3];
    const int mw = m < 9 ? m : 0;                           // columns 9..15 are zero
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float wv5 = w5[(3 * mw + dz) * CH + 8 * kq + e];
            B[dz][e] = (__bf16)(m < 9 ? wv5 : 0.0f);
        }
    }
    const int nzb = ZN - z0 > 16 ? 2 : 1;                   // z blocks of 16 that hold voxels
    const int nit = kIX * nzb;                              // items (input column, z block) per row

    const bf16x8 zero8 = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int kNA = (kBrickP * CH / 8 + 255) / 256;
    bf16x8 pa[kNA];
    // staging index i -> (brick row, chunk), the lane pattern of the operand reads
    auto srow = [](int i) { return ((i >> 6) << 4) | (i & 15); };
    auto schunk = [](int i) { return (i >> 4) & 3; };
    auto fetch = [&](int yy) {
        const bool row_in = (unsigned)yy < (unsigned)YN;
#pragma unroll
        for (int j = 0; j < kNA; ++j) {
            const int i = threadIdx.x + 256 * j;
            const int p = srow(i), q = schunk(i);
            const int gx = x0 - 1 + p / kBZ, gz = z0 - 1 + p % kBZ;
            pa[j] = zero8;
            if (p < kBrick && row_in && (unsigned)gx < (unsigned)XN && (unsigned)gz < (unsigned)ZN)
                pa[j] = ((const bf16x8*)in)[(((long)yy * XN + gx) * ZN + gz) * (CH / 8) + q];
        }
    };
    const float bias = b5[0];
    double r2 = 0.0;
    fetch(y0 - 1);
    // step g: input row gy = y0 - 1 + g gives its shares to output rows gy + 1, gy, gy - 1,
    // then output row gy - 1 is complete
    for (int g = 0; g < ny + 2; ++g) {
        const int gy = y0 - 1 + g;
#pragma unroll
        for (int j = 0; j < kNA; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (srow(i) < kBrick) sA[schunk(i) * kBrickP + srow(i)] = pa[j];
        }
        __syncthreads();   // row staged; the previous step's finalisation is done with its slot
        if (g + 1 < ny + 2) fetch(gy + 1);                  // in flight during this row
        for (int it = wv; it < nit; it += 4) {              // wave-uniform trip count
            const int ix = it / nzb, zb = it - ix * nzb;
            f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int dz = 0; dz < 3; ++dz) {
                // rows z0 + 16 zb + m + dz - 1: brick z 16 zb + m + dz
                const bf16x8 a = sA[kq * kBrickP + ix * kBZ + 16 * zb + m + dz];
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, B[dz], acc, 0, 0, 0);
            }
            // D: lane holds rows z = 16*zb + 4*kq + i of column n = m (dy = m / 3, dx = m % 3):
            // output row gy + 1 - dy, tile column ix - dx
            const int oy = gy + 1 - m / 3 - y0, ox = ix - m % 3;
            if (m < 9 && (unsigned)oy < (unsigned)ny && (unsigned)ox < (unsigned)kLX)
                sD[(((oy + 3) % 3) * kRow + (m * kLX + ox) * kTZ + 16 * zb + 4 * kq) >> 2] = acc;
        }
        __syncthreads();   // shares of this row written; sA free for the next row
        const int oy = gy - 1 - y0;                         // the completed output row
        if (oy >= 0) {
            const float* D = (const float*)sD + (oy % 3) * kRow;
            const int col = threadIdx.x / kTZ, zz = threadIdx.x - col * kTZ;
            const int gx = x0 + col, gz = z0 + zz;
            if (gx < XN && gz < ZN) {
                float sum = bias;
#pragma unroll
                for (int n = 0; n < 9; ++n) sum += D[(n * kLX + col) * kTZ + zz];
                const long v = ((long)(y0 + oy) * XN + gx) * ZN + gz;
                const float o = x[v] + sum;
                out[v] = o;
                if (target) {
                    const float rr = o - target[v];
                    g5[v] = rr * gscale;
                    r2 += (double)rr * rr;
                }
            }
        }
    }
    if (part) {
        for (int sh = 32; sh > 0; sh >>= 1) r2 += __shfl_xor(r2, sh, 64);
        if (lane == 0) sred[wv] = r2;
        __syncthreads();
        if (threadIdx.x == 0) part[tile] = sred[0] + sred[1] + sred[2] + sred[3];
    }
}

// input gradient of the last layer, masked by layer-4 activations:
// g4[v][ci] = (a4[v][ci] > 0) * sum_t w5[t][ci] * g5[v - off(t)]
__global__ void __launch_bounds__(256)
k_dgrad_last(const float* __restrict__ g5, const float* __restrict__ w5, const __bf16* __restrict__ a4,
             __bf16* __restrict__ g_cl, __bf16* __restrict__ g_cm, int XN, int YN, int ZN, int rb)
{
    __shared__ float sW[27 * CH];
    for (int i = threadIdx.x; i < 27 * CH; i += 256) sW[i] = w5[i];
    __syncthreads();
    const long nvox = (long)XN * YN * ZN;
    const long v = blockIdx.x * 256L + threadIdx.x;
    if (v >= nvox) return;
    const int z = (int)(v % ZN);
    const long xy = v / ZN;
    const int xx = (int)(xy % XN), y = (int)(xy / XN);
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0f;
    for (int t = 0; t < 27; ++t) {
        // out[v'] used in[v] through tap t when v' = v - off(t)
        const int tt = 26 - t;   // tap with the opposite offset
        if (!tap_ok(tt, y, xx, z, YN, XN, ZN)) continue;
        const float gv = g5[v + tap_off(tt, XN, ZN)];
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] += sW[t * CH + c] * gv;
    }
    __bf16 o[CH];
    const bf16x8* am = (const bf16x8*)(a4 + v * CH);
#pragma unroll
    for (int q = 0; q < CH / 8; ++q) {
        const bf16x8 mv = am[q];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = 8 * q + e;
            o[c] = (__bf16)((float)mv[e] > 0.0f ? acc[c] : 0.0f);
            g_cm[cm_index(c, v, nvox, rb)] = o[c];
        }
    }
#pragma unroll
    for (int q = 0; q < CH / 8; ++q)
        *(bf16x8*)(g_cl + v * CH + 8 * q) = (bf16x8){o[8 * q], o[8 * q + 1], o[8 * q + 2], o[8 * q + 3],
                                                     o[8 * q + 4], o[8 * q + 5], o[8 * q + 6], o[8 * q + 7]};
}

// ---------------------------------------------------------------------------
// weight gradient of a 32 -> 32 layer on MFMA, K = voxels:
//   dW[t][co][ci] = sum_v g[v][co] * a[v + off(t)][ci]
// from channel-major g_cm, a_cm.  One block = one (dy, dx) pair (3 taps,
// dz = -1, 0, +1) over a chunk of voxels, each of its 4 wavefronts a quarter
// of the chunk (their sums added in LDS in a fixed order); needs ZN % 8 == 0
// so that a run of 8 voxels stays in one z-column.  The dz = +-1 runs are the
// aligned run shifted by one element, assembled with alignbyte from a 16-byte
// load and one neighbouring dword.  Partial sums per chunk:
// part[chunk][27][32][32].
// ---------------------------------------------------------------------------
__device__ __forceinline__ bf16x8 as_bf(u32x4 u)
{
    return __builtin_bit_cast(bf16x8, u);
}

__global__ void __launch_bounds__(256)
k_wgrad32(const __bf16* __restrict__ g_cm, const __bf16* __restrict__ a_cm, float* __restrict__ part,
          int XN, int YN, int ZN, long chunk)
{
    __shared__ float sR[3][48][64];               // 36,864 B: wavefronts 1..3's sums
    const long nvox = (long)XN * YN * ZN;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const int pair = blockIdx.y;                   // 0..8 = (dy+1)*3 + (dx+1)
    const int dy = pair / 3 - 1, dx = pair % 3 - 1;
    const long sub = chunk / 4;                    // chunk % 128 == 0
    const long c0 = blockIdx.x * chunk + wv * sub;
    const long c1 = min(c0 + sub, nvox);
    f32x4 acc[3][2][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[d][i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const long shift = (long)dx * ZN + (long)dy * XN * ZN;
    const uint32_t* a32 = (const uint32_t*)a_cm;
    // operands of one 32-voxel step; the next step's are loaded while the
    // current step's 12 MFMAs run.  This lane's run of 8 voxels starts at
    // vr = v0 + 8*kq, whose (z, x, y) are stepped, not divided (ZN % 8 == 0:
    // a run never crosses a column)
    struct Ops {
        bf16x8 A0, A1;
        u32x4 blk[2];
        uint32_t prev[2], next[2];
        int z0;
    };
    long vr = c0 + 8 * kq;
    int rz, rx, ry;
    {
        const long vc = vr < nvox ? vr : 0;
        rz = (int)(vc % ZN);
        const long xy = vc / ZN;
        rx = (int)(xy % XN);
        ry = (int)(xy / XN);
    }
    auto load = [&](Ops& o) {
        const bool in = vr < nvox;
        const bool ok = in && (unsigned)(ry + dy) < (unsigned)YN && (unsigned)(rx + dx) < (unsigned)XN;
        o.z0 = rz;
        o.A0 = o.A1 = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
        if (in) {
            o.A0 = *(const bf16x8*)(g_cm + (long)m * nvox + vr);
            o.A1 = *(const bf16x8*)(g_cm + (long)(16 + m) * nvox + vr);
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            o.blk[nb] = (u32x4){0u, 0u, 0u, 0u};
            o.prev[nb] = o.next[nb] = 0u;
            if (ok) {
                const long e = (long)(nb * 16 + m) * nvox + vr + shift;     // 8-aligned element index
                o.blk[nb] = *(const u32x4*)(a_cm + e);
                const long ep = e - 2 >= 0 ? e - 2 : 0;
                const long en = e + 8 <= (long)CH * nvox - 2 ? e + 8 : e;
                o.prev[nb] = a32[ep >> 1];
                o.next[nb] = a32[en >> 1];
            }
        }
        // advance the run by 32 voxels
        vr += 32;
        rz += 32;
        while (rz >= ZN) {
            rz -= ZN;
            if (++rx == XN) { rx = 0; ++ry; }
        }
    };
    Ops cur, nxt;
    load(cur);
    for (long v0 = c0; v0 < c1; v0 += 32) {
        if (v0 + 32 < c1) load(nxt);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const u32x4 blk = cur.blk[nb];
            // dz = 0
            const bf16x8 B0 = as_bf(blk);
            // dz = -1: elements e-1 .. e+6 (element 0 invalid at z0 == 0)
            u32x4 lo;
            lo[0] = __builtin_amdgcn_alignbyte(blk[0], cur.prev[nb], 2);
            lo[1] = __builtin_amdgcn_alignbyte(blk[1], blk[0], 2);
            lo[2] = __builtin_amdgcn_alignbyte(blk[2], blk[1], 2);
            lo[3] = __builtin_amdgcn_alignbyte(blk[3], blk[2], 2);
            if (cur.z0 == 0) lo[0] &= 0xffff0000u;
            // dz = +1: elements e+1 .. e+8 (element 7 invalid at z0 + 8 == ZN)
            u32x4 hi;
            hi[0] = __builtin_amdgcn_alignbyte(blk[1], blk[0], 2);
            hi[1] = __builtin_amdgcn_alignbyte(blk[2], blk[1], 2);
            hi[2] = __builtin_amdgcn_alignbyte(blk[3], blk[2], 2);
            hi[3] = __builtin_amdgcn_alignbyte(cur.next[nb], blk[3], 2);
            if (cur.z0 + 8 == ZN) hi[3] &= 0x0000ffffu;
            const bf16x8 Bm = as_bf(lo), Bp = as_bf(hi);
            acc[0][0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A0, Bm, acc[0][0][nb], 0, 0, 0);
            acc[0][1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A1, Bm, acc[0][1][nb], 0, 0, 0);
            acc[1][0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A0, B0, acc[1][0][nb], 0, 0, 0);
            acc[1][1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A1, B0, acc[1][1][nb], 0, 0, 0);
            acc[2][0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A0, Bp, acc[2][0][nb], 0, 0, 0);
            acc[2][1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.A1, Bp, acc[2][1][nb], 0, 0, 0);
        }
        cur = nxt;
    }
    // wavefronts 1..3 hand their sums to wavefront 0, which adds them in order
    if (wv > 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) sR[wv - 1][((d * 2 + mb) * 2 + nb) * 4 + i][lane] = acc[d][mb][nb][i];
    }
    __syncthreads();
    if (wv > 0) return;
#pragma unroll
    for (int w = 0; w < 3; ++w)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[d][mb][nb][i] += sR[w][((d * 2 + mb) * 2 + nb) * 4 + i][lane];
    // D: rows co = mb*16 + 4*kq + i, column ci = nb*16 + m
    float* P = part + (long)blockIdx.x * 27 * CH * CH;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int t = pair * 3 + d;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    P[((long)t * CH + mb * 16 + 4 * kq + i) * CH + nb * 16 + m] = acc[d][mb][nb][i];
    }
}

// ---------------------------------------------------------------------------
// Weight and bias gradients of a 32 -> 32 layer on MFMA, swept along x
// (ZN % 32 == 0; g and a as run-blocked copies):
//   dW[t][co][ci] = sum_v g[v][co] a[v + off(t)][ci],   db[co] = sum_v g[v][co]
// With z' = z + dz the z shift of tap t moves onto g:
//   dW[t][co][ci] = sum_v' g[v' - dz][co] a[v' + off_xy(t)][ci]
// (g and a read as 0 outside the volume), so each (dy, dx) column of a enters
// its MFMAs unshifted and only g needs the dz = +-1 runs (alignbyte of its
// 16-byte load and two neighbouring dwords).  Wavefront w = (cb, nb) of a
// block owns the 16 x 16 block co in [16cb, 16cb+16), ci in [16nb, 16nb+16)
// of all 27 taps (27 accumulators) and walks runs of 32 z (the K of one MFMA)
// along x, holding the a columns at x-1, x, x+1 of rows y-1, y, y+1 in a ring
// of kSwR slots: per step it loads g at x + kSwD and the three a columns at
// x + 1 + kSwD (kSwD steps ahead) and issues 27 MFMAs and one for the bias
// (B = 1): 6 loads per 28 MFMAs, each load one contiguous KB of a run block.
// (Channel-major rows 2 * nvox bytes apart made each load touch 16 lines and
// pages: 255 us instead of 137 us at 256 x 256 x 32.)
// Work = the x-steps of all lines (y, z-run), dealt to blocks as contiguous
// ranges, the ranges of one XCD contiguous (neighbouring lines read the same
// rows of a through that XCD's L2).  Reads go through buffer resources whose
// out-of-range offset reads 0: that supplies the zeros outside the volume and
// past the end of a segment (steps past it multiply g = 0), so the unrolled
// steps carry no branches.
// Partials: part[block][27*32*32 + 32] = the layer's W gradient, then b.
// ---------------------------------------------------------------------------
constexpr int kSwD = 4;                 // prefetch distance (steps)
constexpr int kSwR = kSwD + 3;          // ring slots (= steps per unrolled group)
constexpr unsigned kBias = 0x60000000u; // added to an offset: past every resource (reads 0)
constexpr int kBufW3 = 0x00020000;      // buffer-resource word 3: raw access on gfx9

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_wgrad32_sweep(const __bf16* __restrict__ g_rb, const __bf16* __restrict__ a_rb, float* __restrict__ part,
                int XN, int YN, int ZN)
{
    const long nvox = (long)XN * YN * ZN;
    const int nr = ZN >> 5;
    const long T = (long)YN * nr * XN;             // x-steps of all lines
    const long NB = gridDim.x;                     // a multiple of 8
    const long j = (long)(blockIdx.x & 7) * (NB >> 3) + (blockIdx.x >> 3);
    const long s0 = j * T / NB, s1 = (j + 1) * T / NB;

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: scalar resources
    const int m = lane & 15, kq = lane >> 4;
    const int cb = wv >> 1, nb = wv & 1;
    // one resource per tensor, nbytes = 64 * nvox <= 2^30 (host check): an
    // offset holding one or two kBias stays below 2^32 and past the resource
    const unsigned nbytes = (unsigned)(64 * nvox);
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc((void*)g_rb, (short)0, (int)nbytes, kBufW3);
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc((void*)a_rb, (short)0, (int)nbytes, kBufW3);
    // this lane's 8 elements in a run block: channel 16cb + m (g) / 16nb + m (a), z 8kq..8kq+7
    const unsigned lane_g = (unsigned)(((cb * 16 + m) * 32 + 8 * kq) * 2);
    const unsigned lane_a = (unsigned)(((nb * 16 + m) * 32 + 8 * kq) * 2);

    f32x4 acc[27], accb = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const bf16x8 ones = __builtin_bit_cast(bf16x8, (u32x4){0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u});

    for (long s = s0; s < s1;) {                   // one segment of a line per pass (wave-uniform)
        const long ln = s / XN;
        const int x0 = (int)(s - ln * XN);
        const int xend = (int)min((long)XN, x0 + (s1 - s));
        const int y = (int)(ln / nr), r = (int)(ln - (long)y * nr);
        // byte offsets: lane part + uniform part (scalar); an invalid row,
        // column or step adds kBias, which takes the sum past the resource.
        // Run (y, x, r) starts at byte ((y * XN + x) * nr + r) * 2048.
        unsigned rowb[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int yy = y + d - 1;
            rowb[d] = (unsigned)yy < (unsigned)YN ? (unsigned)(((long)yy * XN * nr + r) * 2048) : kBias;
        }
        const unsigned gbase = (unsigned)(((long)y * XN * nr + r) * 2048);
        const unsigned cstep = (unsigned)nr * 2048u;
        const int cmax = min(XN, xend + 1);               // columns x0-1 .. xend are read
        // g at z - 2, z - 1 and z + 8, z + 9 of the lane's 8: inside its own
        // channel row, in the previous / next run of the column (+-1024
        // elements), or outside the column (kBias)
        const unsigned padj = kq > 0 ? (unsigned)-4 : r > 0 ? (unsigned)(-(1024 - 30) * 2) : kBias;
        const unsigned nadj = kq < 3 ? 16u : r < nr - 1 ? (unsigned)((1024 - 24) * 2) : kBias;

        u32x4 col[kSwR][3];                        // a columns, slot (c - x0 + 1) % kSwR
        u32x4 gb[kSwR];                            // g at x, slot (x - x0) % kSwR
        uint32_t gp[kSwR], gn[kSwR];               // its neighbours z-1 (elements -2, -1), z+8 (8, 9)
        auto load_col = [&](u32x4 (&c3)[3], int c) {
            const unsigned cofs = (unsigned)c < (unsigned)cmax ? (unsigned)c * cstep : kBias;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                c3[d] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(
                                                      ra, lane_a + (rowb[d] + cofs), 0, 0));
        };
        auto load_g = [&](u32x4& b, uint32_t& p, uint32_t& n, int x) {
            const unsigned off = lane_g + (x < xend ? gbase + (unsigned)x * cstep : kBias);
            b = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rg, off, 0, 0));
            p = __builtin_amdgcn_raw_buffer_load_b32(rg, off + padj, 0, 0);
            n = __builtin_amdgcn_raw_buffer_load_b32(rg, off + nadj, 0, 0);
        };
#pragma unroll
        for (int i = 0; i < kSwD + 2; ++i) load_col(col[i], x0 - 1 + i);
#pragma unroll
        for (int i = 0; i < kSwD; ++i) load_g(gb[i], gp[i], gn[i], x0 + i);

        for (int base = x0; base < xend; base += kSwR) {
#pragma unroll
            for (int k = 0; k < kSwR; ++k) {
                const int x = base + k;
                // keep each step's operand math after its loads: hoisted next to
                // them, it would wait on the loads just issued
                __builtin_amdgcn_sched_barrier(0);
                load_col(col[(k + kSwR - 1) % kSwR], x + 1 + kSwD);
                load_g(gb[(k + kSwD) % kSwR], gp[(k + kSwD) % kSwR], gn[(k + kSwD) % kSwR], x + kSwD);
                const u32x4 b = gb[k];
                // dz = +1 pairs a at z' with g at z' - 1: elements -1 .. 6 of the run
                u32x4 lo;
                lo[0] = __builtin_amdgcn_alignbyte(b[0], gp[k], 2);
                lo[1] = __builtin_amdgcn_alignbyte(b[1], b[0], 2);
                lo[2] = __builtin_amdgcn_alignbyte(b[2], b[1], 2);
                lo[3] = __builtin_amdgcn_alignbyte(b[3], b[2], 2);
                // dz = -1 pairs a at z' with g at z' + 1: elements 1 .. 8
                u32x4 hi;
                hi[0] = __builtin_amdgcn_alignbyte(b[1], b[0], 2);
                hi[1] = __builtin_amdgcn_alignbyte(b[2], b[1], 2);
                hi[2] = __builtin_amdgcn_alignbyte(b[3], b[2], 2);
                hi[3] = __builtin_amdgcn_alignbyte(gn[k], b[3], 2);
                const bf16x8 Gm = as_bf(hi), G0 = as_bf(b), Gp = as_bf(lo);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        const bf16x8 B = as_bf(col[(k + dx) % kSwR][dy]);
                        const int t = dy * 9 + dx * 3;
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Gm, B, acc[t], 0, 0, 0);
                        acc[t + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(G0, B, acc[t + 1], 0, 0, 0);
                        acc[t + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Gp, B, acc[t + 2], 0, 0, 0);
                    }
                accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(G0, ones, accb, 0, 0, 0);
            }
        }
        s += xend - x0;
    }
    // D: rows co = 16cb + 4kq + i, column ci = 16nb + m
    float* P = part + j * (27 * CH * CH + CH);
#pragma unroll
    for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) P[((long)t * CH + cb * 16 + 4 * kq + i) * CH + nb * 16 + m] = acc[t][i];
    if (nb == 0 && m == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) P[27 * CH * CH + cb * 16 + 4 * kq + i] = accb[i];
    }
}

// Weight gradients of the thin layers (VALU), one block per chunk of voxels,
// partials part[chunk][27*32].  Thread r = tid % 128 < 108 owns tap t = r / 4
// and the 8 channels 8*(r % 4)..+7 of it: per voxel one 16-byte load of those
// channels and one scalar feed 8 FMAs.  The block's two halves (tid / 128)
// take the two halves of the chunk, 4 voxels per step (chunk and ZN are
// multiples of 8, so the 4 share a z-column; coordinates are stepped), and
// their sums are added in LDS in a fixed order.  Loads outside the volume go
// to a valid index and their value is replaced by 0, so none waits on a branch.
struct ThinTap {
    int t, o, dy, dx, dz;
    long off;
    bool live;
};
__device__ __forceinline__ ThinTap thin_tap(int XN, int ZN)
{
    ThinTap q;
    const int r = threadIdx.x & 127;
    q.live = r < 27 * 4;
    q.t = q.live ? r >> 2 : 0;
    q.o = r & 3;
    q.dy = q.t / 9 - 1;
    q.dx = (q.t / 3) % 3 - 1;
    q.dz = q.t % 3 - 1;
    q.off = tap_off(q.t, XN, ZN);
    return q;
}

// the two halves' sums -> part[blockIdx.x][t*32 + 8*o + e]
__device__ __forceinline__ void thin_store(const ThinTap& q, float (&acc)[8], float (*sR)[8], float* part)
{
    const int r = threadIdx.x & 127;
    if (threadIdx.x >= 128) {
#pragma unroll
        for (int e = 0; e < 8; ++e) sR[r][e] = acc[e];
    }
    __syncthreads();
    if (threadIdx.x < 128 && q.live) {
#pragma unroll
        for (int e = 0; e < 8; ++e) part[(long)blockIdx.x * 27 * CH + q.t * CH + 8 * q.o + e] = acc[e] + sR[r][e];
    }
}

// weight gradient of the first layer: dW1[t][co] = sum_v g1[v][co] * x[v + off(t)]
// (g1 channels-last)
__global__ void __launch_bounds__(256)
k_wgrad_first(const __bf16* __restrict__ g_cl, const float* __restrict__ x, float* __restrict__ part,
              int XN, int YN, int ZN, long chunk)
{
    __shared__ float sR[128][8];
    const long nvox = (long)XN * YN * ZN;
    const ThinTap q = thin_tap(XN, ZN);
    const long half = chunk / 2;
    const long c0 = blockIdx.x * chunk + (threadIdx.x >> 7) * half, c1 = min(c0 + half, nvox);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int z = 0, xx = 0, y = 0;
    if (c0 < nvox) {
        z = (int)(c0 % ZN);
        const long xy = c0 / ZN;
        xx = (int)(xy % XN);
        y = (int)(xy / XN);
    }
    for (long v = c0; v < c1; v += 4) {
        const bool col = q.live && (unsigned)(y + q.dy) < (unsigned)YN && (unsigned)(xx + q.dx) < (unsigned)XN;
        bf16x8 g[4];
        float xa[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bool ok = col && (unsigned)(z + i + q.dz) < (unsigned)ZN;
            g[i] = *(const bf16x8*)(g_cl + (v + i) * CH + 8 * q.o);
            const float xv = x[ok ? v + i + q.off : v + i];
            xa[i] = ok ? xv : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] += bf(g[i][e]) * xa[i];
        z += 4;
        if (z == ZN) {
            z = 0;
            if (++xx == XN) { xx = 0; ++y; }
        }
    }
    thin_store(q, acc, sR, part);
}

// weight gradient of the last layer: dW5[t][ci] = sum_v g5[v] * a4[v + off(t)][ci]
__global__ void __launch_bounds__(256)
k_wgrad_last(const float* __restrict__ g5, const __bf16* __restrict__ a_cl, float* __restrict__ part,
             int XN, int YN, int ZN, long chunk)
{
    __shared__ float sR[128][8];
    const long nvox = (long)XN * YN * ZN;
    const ThinTap q = thin_tap(XN, ZN);
    const long half = chunk / 2;
    const long c0 = blockIdx.x * chunk + (threadIdx.x >> 7) * half, c1 = min(c0 + half, nvox);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int z = 0, xx = 0, y = 0;
    if (c0 < nvox) {
        z = (int)(c0 % ZN);
        const long xy = c0 / ZN;
        xx = (int)(xy % XN);
        y = (int)(xy / XN);
    }
    for (long v = c0; v < c1; v += 4) {
        const bool col = q.live && (unsigned)(y + q.dy) < (unsigned)YN && (unsigned)(xx + q.dx) < (unsigned)XN;
        bf16x8 a[4];
        float gv[4];
        bool ok[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ok[i] = col && (unsigned)(z + i + q.dz) < (unsigned)ZN;
            gv[i] = g5[v + i];
            a[i] = *(const bf16x8*)(a_cl + (ok[i] ? v + i + q.off : v + i) * CH + 8 * q.o);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float gi = ok[i] ? gv[i] : 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] += gi * bf(a[i][e]);
        }
        z += 4;
        if (z == ZN) {
            z = 0;
            if (++xx == XN) { xx = 0; ++y; }
        }
    }
    thin_store(q, acc, sR, part);
}

// bias-gradient partials: part[j][c] = sum over chunk j of g_cm[c][v] (block j;
// wavefront w takes channels w, w+4, ...; lanes read consecutive voxels).  The
// chunks are then summed in a fixed order (k_reduce_chunks), so the result
// does not depend on scheduling.
__global__ void __launch_bounds__(256)
k_bias_grad_partial(const __bf16* __restrict__ g_cm, long nvox, long chunk, float* __restrict__ part, int rb)
{
    const long c0 = blockIdx.x * chunk, c1 = min(c0 + chunk, nvox);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = wv; c < CH; c += 4) {
        float s = 0.0f;
        for (long v = c0 + lane; v < c1; v += 64) s += bf(g_cm[cm_index(c, v, nvox, rb)]);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) part[(long)blockIdx.x * CH + c] = s;
    }
}

// sum partials: part[j] = sum over chunk j of x[v]
__global__ void __launch_bounds__(256)
k_sum_partial(const float* __restrict__ x, long n, long chunk, float* __restrict__ part)
{
    __shared__ float sh[4];
    const long c0 = blockIdx.x * chunk, c1 = min(c0 + chunk, n);
    float s = 0.0f;
    for (long v = c0 + threadIdx.x; v < c1; v += 256) s += x[v];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// out[y*n + i] = sum of part[c*n + i] over the chunks c of group y
// ([y*per, min((y+1)*per, nchunks))).  Block = 64 outputs (lanes) x 4
// wavefronts, each summing one quarter of the group's chunks in order; the
// quarters are added in a fixed order, so the result does not depend on
// scheduling.
__global__ void __launch_bounds__(256)
k_reduce_chunks(const float* __restrict__ part, int nchunks, int per, int n, float* __restrict__ out)
{
    __shared__ float s4[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const int c0 = blockIdx.y * per, c1 = min(c0 + per, nchunks);
    const int q = (c1 - c0 + 3) / 4;
    const int a = c0 + w * q, b = min(a + q, c1);
    float s = 0.0f;
    if (i < n)
        for (int c = a; c < b; ++c) s += part[(long)c * n + i];
    s4[w][lane] = s;
    __syncthreads();
    if (w == 0 && i < n) out[(long)blockIdx.y * n + i] = (s4[0][lane] + s4[1][lane]) + (s4[2][lane] + s4[3][lane]);
}

// flipped, transposed weights for the input gradient: wt[t][ci][co] = w[26-t][co][ci]
__global__ void k_flip_transpose(const __bf16* __restrict__ w, __bf16* __restrict__ wt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 27 * CH * CH) return;
    const int t = i / (CH * CH), ci = (i / CH) % CH, co = i % CH;
    wt[i] = w[((26 - t) * CH + co) * CH + ci];
}

// Adam on FP32 master parameters; also refreshes the BF16 copies the kernels use
// (flat parameter vector, per-layer BF16 images written by index maps)
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// bf16 image of an fp32 parameter block (rounded to nearest even), and an fp32
// copy holding the same rounded values for the VALU layers
__global__ void k_to_bf16(const float* __restrict__ src, __bf16* __restrict__ dst_bf, float* __restrict__ dst_f,
                          long n)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const __bf16 b = (__bf16)src[i];
    if (dst_bf) dst_bf[i] = b;
    if (dst_f) dst_f[i] = (float)b;
}

}  // namespace ctlab

// Kernel-predicting U-Net denoiser on gfx950 bf16 MFMA (see sr_cnn.h for the topology).
//
// A 3x3 convolution is an implicit GEMM: M = output pixels, N = output channels, K = 9 taps x
// input channels. A workgroup (4 waves) owns a 16 x 8 pixel tile; each wave owns two tile rows
// (two 16-pixel M tiles) and every output channel (NT tiles of 16). Per 32-channel K chunk the
// workgroup stages the 18 x 10 halo tile in LDS (fusing 2x2 average pooling or nearest
// upsampling + skip concatenation into the staging), then each of the 9 taps reads one A fragment
// per M tile from LDS and one B fragment per N tile (16 contiguous bytes of [tap][cout][cin]
// weights, shared by all workgroups through the caches) and issues v_mfma_f32_16x16x32_bf16.
// Operand layouts of that instruction (checked on the hardware): A[i][k] sits in lane
// i + 16 (k / 8), element k % 8; B[k][n] in lane n + 16 (k / 8), element k % 8; D[i][n] in lane
// n + 16 (i / 4), register i % 4.
#include "sr_cnn.h"

#include <math.h>
#include <string.h>

#include <algorithm>

#define SR_CNN_DEV __device__ __forceinline__

namespace sr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTW = 16, kTH = 8;               // output tile
constexpr int kHW = kTW + 2, kHH = kTH + 2;    // halo tile
constexpr int kHalo = kHW * kHH;               // 180 pixels
constexpr int kFeat = 10;                      // input feature channels
constexpr int kTaps = 25;                      // 5 x 5 predicted kernel
constexpr int kStages = kKpnStages;            // kernel cascade, taps 1, 2, 4 and 8 pixels apart
constexpr int kLogits = kTaps * kStages;       // 100 logits per pixel
constexpr int kLogitStride = 128;              // their padded channel stride
constexpr int kStageStride = 32;               // stage s's logits sit in channels [32 s, 32 s + 25)

int pad32(int c) { return (c + 31) & ~31; }

SR_CNN_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
SR_CNN_DEV uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
SR_CNN_DEV uint32_t pack2(float a, float b) { return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16); }

struct ConvArgs {
    const uint16_t* a;  // source (mode 2: the low-resolution input that is upsampled)
    const uint16_t* b;  // skip source (mode 2)
    int ca, cb;         // padded channel strides of a and b
    int wa, ha;         // dimensions of a
    int mode;           // 0 same resolution, 1 2x2 average pool of a, 2 concat(up(a), b)
    int W, H;           // output (and b) dimensions
    int cinPad, coutPad, cout;
    const uint16_t* w;  // [9][coutPad][cinPad]
    const float* bias;  // [coutPad]
    uint16_t* out;      // [pixel][coutPad], or with plane > 0 [coutPad / 32][pixel][32]
    size_t plane;       // elements between those 32-channel planes (0: interleaved output)
    int relu;
};

// Address of output channels [c, c + 4) of pixel `pixel` (c a multiple of 4).
SR_CNN_DEV uint16_t* out_at(const ConvArgs& A, size_t pixel, int c) {
    if (A.plane) return A.out + (size_t)(c >> 5) * A.plane + pixel * 32 + (c & 31);
    return A.out + pixel * A.coutPad + c;
}

// 8 consecutive input channels (16 bytes) of K chunk c0 at output-resolution pixel (y, x), zero
// outside the image.
SR_CNN_DEV uint4 fetch8(const ConvArgs& A, int c, int y, int x) {
    uint4 z = make_uint4(0, 0, 0, 0);
    if (x < 0 || y < 0 || x >= A.W || y >= A.H) return z;
    if (A.mode == 0) return *(const uint4*)(A.a + ((size_t)y * A.W + x) * A.ca + c);
    if (A.mode == 2) {
        if (c < A.ca) return *(const uint4*)(A.a + ((size_t)(y >> 1) * A.wa + (x >> 1)) * A.ca + c);
        return *(const uint4*)(A.b + ((size_t)y * A.W + x) * A.cb + (c - A.ca));
    }
    // 2x2 average of the finer source, the last row / column repeated at odd sizes.
    const int x0 = 2 * x, y0 = 2 * y;
    const int x1 = x0 + 1 < A.wa ? x0 + 1 : A.wa - 1, y1 = y0 + 1 < A.ha ? y0 + 1 : A.ha - 1;
    const uint4 p00 = *(const uint4*)(A.a + ((size_t)y0 * A.wa + x0) * A.ca + c);
    const uint4 p01 = *(const uint4*)(A.a + ((size_t)y0 * A.wa + x1) * A.ca + c);
    const uint4 p10 = *(const uint4*)(A.a + ((size_t)y1 * A.wa + x0) * A.ca + c);
    const uint4 p11 = *(const uint4*)(A.a + ((size_t)y1 * A.wa + x1) * A.ca + c);
    const uint32_t* q00 = (const uint32_t*)&p00;
    const uint32_t* q01 = (const uint32_t*)&p01;
    const uint32_t* q10 = (const uint32_t*)&p10;
    const uint32_t* q11 = (const uint32_t*)&p11;
    uint32_t r[4];
    for (int i = 0; i < 4; ++i) {
        const float lo = ((bf2f((uint16_t)q00[i]) + bf2f((uint16_t)q01[i])) + bf2f((uint16_t)q10[i])) +
                         bf2f((uint16_t)q11[i]);
        const float hi = ((bf2f((uint16_t)(q00[i] >> 16)) + bf2f((uint16_t)(q01[i] >> 16))) +
                          bf2f((uint16_t)(q10[i] >> 16))) + bf2f((uint16_t)(q11[i] >> 16));
        r[i] = pack2(lo * 0.25f, hi * 0.25f);
    }
    return make_uint4(r[0], r[1], r[2], r[3]);
}

// LDS slot of 16-byte group `part` of row p (a halo pixel's 32 channels, or a weight row's 32
// input channels): the groups of row p are rotated by (p >> 1) & 3, so that the 8 lanes of a
// 128-byte LDS cycle, reading one group of 8 consecutive rows, hit 8 different bank groups.
SR_CNN_DEV int lds_slot(int p, int part) { return p * 4 + (part ^ ((p >> 1) & 3)); }

// Epilogue of a wave's MT rows x NT channel tiles from MFMAs with the weights as A: acc[mt][nt][r]
// is channel c0 + 16 nt + 4 grp + r of pixel (x, y0 + mt). Bias, ReLU and bf16 rounding, then one
// 8-byte store per 4 channels; channels at or past cout are written as 0.
template <int NT, int MT>
SR_CNN_DEV void store_tile(const ConvArgs& A, const f32x4 (&acc)[MT][NT], int x, int y0, int c0, int grp) {
    float bias[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = c0 + nt * 16 + grp * 4 + r;
            bias[nt][r] = co < A.cout ? A.bias[co] : 0.0f;
        }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int y = y0 + mt;
        if (x >= A.W || y >= A.H) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[mt][nt][r] + bias[nt][r];
                if (A.relu) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
                if (c0 + nt * 16 + grp * 4 + r >= A.cout) v[r] = 0.0f;
            }
            *(uint2*)out_at(A, (size_t)y * A.W + x, c0 + nt * 16 + grp * 4) =
                make_uint2(pack2(v[0], v[1]), pack2(v[2], v[3]));
        }
    }
}

// A 4-wave workgroup computes a 16-wide, 4*MT-tall output tile: wave w owns rows w*MT .. w*MT+MT-1,
// one 16-pixel MFMA M tile per row. Per 32-channel K chunk the (4*MT+2) x 18 input halo is staged
// in LDS; per tap each wave loads its NT weight fragments once and reuses them for its MT rows.
// The result is transposed through LDS so that every lane stores whole 16-byte groups of a pixel's
// channels.
// Waves per SIMD the register allocation of k_conv3x3<NT, MT> must allow (1: the compiler's choice).
#ifndef SR_W2_4
#define SR_W2_4 3
#endif
template <int NT, int MT>
constexpr int conv_waves() { return NT == 2 && MT == 4 ? SR_W2_4 : 1; }
// k_conv3x3<NT, MT> stages its halo with all of a thread's loads in flight at once (else one
// load-and-store round trip at a time, which takes fewer registers).
#ifndef SR_BATCH2_4
#define SR_BATCH2_4 1
#endif
template <int NT, int MT>
constexpr bool conv_batched() { return NT == 2 && MT == 4 && SR_BATCH2_4; }

// Halo words [0, N) of K chunk c0 of a MODE source into LDS slots with NT threads, all of a
// thread's loads issued before its first LDS store.
template <int MODE, int N, int NT>
SR_CNN_DEV void stage_halo(const ConvArgs& A, int c0, int tx0, int ty0, uint4* tile) {
    constexpr int kIt = (N + NT - 1) / NT;
    uint4 v[kIt];
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
        const int idx = min((int)threadIdx.x + NT * i, N - 1);
        const int p = idx >> 2, part = idx & 3, y = ty0 - 1 + p / kHW, x = tx0 - 1 + p % kHW;
        const bool ok = x >= 0 && y >= 0 && x < A.W && y < A.H;
        const int yc = ok ? y : 0, xc = ok ? x : 0, c = c0 + part * 8;
        uint4 w;
        if (MODE == 0) w = *(const uint4*)(A.a + ((size_t)yc * A.W + xc) * A.ca + c);
        else w = c < A.ca ? *(const uint4*)(A.a + ((size_t)(yc >> 1) * A.wa + (xc >> 1)) * A.ca + c)
                          : *(const uint4*)(A.b + ((size_t)yc * A.W + xc) * A.cb + (c - A.ca));
        v[i] = ok ? w : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
        const int idx = (int)threadIdx.x + NT * i;
        if (idx < N) tile[lds_slot(idx >> 2, idx & 3)] = v[i];
    }
}

template <int NT, int MT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(conv_waves<NT, MT>())))
k_conv3x3(ConvArgs A) {
    constexpr int kRows = 4 * MT;
    constexpr int kHaloPx = kHW * (kRows + 2);
    __shared__ uint4 tile[kHaloPx * 4];  // halo pixels x 32 channels of bf16, rows as lds_slot
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx0 = blockIdx.x * kTW, ty0 = blockIdx.y * kRows;
    f32x4 acc[MT][NT];
    #pragma unroll
    for (int mt = 0; mt < MT; ++mt)
        #pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < A.cinPad; c0 += 32) {
        __syncthreads();
        if (conv_batched<NT, MT>() && A.mode == 0) stage_halo<0, kHaloPx * 4, 256>(A, c0, tx0, ty0, tile);
        else if (conv_batched<NT, MT>() && A.mode == 2) stage_halo<2, kHaloPx * 4, 256>(A, c0, tx0, ty0, tile);
        else
            for (int idx = threadIdx.x; idx < kHaloPx * 4; idx += 256) {
                const int p = idx >> 2, part = idx & 3;
                tile[lds_slot(p, part)] = fetch8(A, c0 + part * 8, ty0 - 1 + p / kHW, tx0 - 1 + p % kHW);
            }
        __syncthreads();
        const uint16_t* wc = A.w + c0 + (lane >> 4) * 8;
        #pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            const uint16_t* wt = wc + (size_t)tap * A.coutPad * A.cinPad;
            bf16x8 wfr[NT];
            #pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                wfr[nt] = __builtin_bit_cast(bf16x8, *(const uint4*)(wt + (size_t)(nt * 16 + (lane & 15)) * A.cinPad));
            #pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int p = (wave * MT + mt + dy) * kHW + (lane & 15) + dx;
                const bf16x8 px = __builtin_bit_cast(bf16x8, tile[lds_slot(p, lane >> 4)]);
                #pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfr[nt], px, acc[mt][nt], 0, 0, 0);
            }
        }
    }
    // acc[mt][nt][r]: channel 16 nt + 4 (lane >> 4) + r of pixel (tx0 + (lane & 15), row wave MT + mt):
    // bias, ReLU and bf16 rounding, stored as one 8-byte word per 4 channels (channels past cout as 0).
    store_tile<NT>(A, acc, tx0 + (lane & 15), ty0 + wave * MT, 0, lane >> 4);
}

// Layers split by output channels: wave w of the NW-wave workgroup (NW = coutPad / 32) computes
// the whole 16 x MT tile for output channels [32 w, 32 w + 32) (two MFMA N tiles), so each wave
// loads only its own weight fragments and the block's pixels are staged once.
template <int MT, int NW>
__global__ void __launch_bounds__(NW * 64) k_conv3x3_split(ConvArgs A) {
    constexpr int kHaloPx = kHW * (MT + 2);
    constexpr int kHaloWords = kHaloPx * 4;
    __shared__ uint4 smem[kHaloWords];  // rows as lds_slot
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx0 = blockIdx.x * kTW, ty0 = blockIdx.y * MT;
    f32x4 acc[MT][2];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt][0] = acc[mt][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < A.cinPad; c0 += 32) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < kHaloWords; idx += NW * 64) {
            const int p = idx >> 2, part = idx & 3;
            smem[lds_slot(p, part)] = fetch8(A, c0 + part * 8, ty0 - 1 + p / kHW, tx0 - 1 + p % kHW);
        }
        __syncthreads();
        const uint16_t* wc = A.w + c0 + (lane >> 4) * 8;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            const uint16_t* wt = wc + (size_t)tap * A.coutPad * A.cinPad;
            bf16x8 wfr[2];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                wfr[h] = __builtin_bit_cast(
                    bf16x8, *(const uint4*)(wt + (size_t)((2 * wave + h) * 16 + (lane & 15)) * A.cinPad));
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int p = (mt + dy) * kHW + (lane & 15) + dx;
                const bf16x8 px = __builtin_bit_cast(bf16x8, smem[lds_slot(p, lane >> 4)]);
                acc[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfr[0], px, acc[mt][0], 0, 0, 0);
                acc[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfr[1], px, acc[mt][1], 0, 0, 0);
            }
        }
    }
    store_tile<2>(A, acc, tx0 + (lane & 15), ty0, 32 * wave, lane >> 4);
}

// Persistent 3x3 convolution. Wave w owns output channels [16 NTW cg, 16 NTW (cg + 1)) with
// cg = w % NG and tile rows [MT rg, MT rg + MT) with rg = w / NG, so a workgroup covers the
// kCo = NG 16 NTW channels [0, kCo) of a 16 x (4 / NG) MT pixel tile, and walks tiles
// t = blockIdx.x, + gridDim.x, ..., each as cinPad / 32 steps of one 32-channel K chunk.
// - Everything a step reads from memory is loaded into registers during the step before: the
//   halo of its K chunk and, with more than one chunk, the chunk's weights (all 9 taps, the
//   block's kCo channels). Both are then written to LDS, shared by the 4 waves. With one chunk each
//   wave keeps its weight fragments in registers for the whole kernel.
// - LDS rows of four 16-byte groups (a halo pixel's 32 channels, or a weight row's 32 input
//   channels) have the groups of row p rotated by (p >> 1) & 3, so that the 8 lanes of a 128-byte
//   LDS cycle read 8 different bank groups.
// - Each pixel fragment is read once per halo row and column offset and used for the (up to)
//   three output rows it feeds; every output still sums its taps in order 0..8 within a chunk
//   and its chunks in order, as k_conv3x3 does.
// - The MFMA takes the weights as A (rows: output channels) and the pixels as B, so each lane ends
//   with 4 consecutive channels of one pixel and stores them as one 8-byte word: no LDS transpose.
// MODE 0 reads a same-resolution source, MODE 2 the concatenation of the nearest-upsampled source
// a (chunks below ca) and the skip source b.
#ifndef SR_PIPE_WAVES
#define SR_PIPE_WAVES 2  // waves per SIMD the register allocation must allow
#endif

template <int MODE>
SR_CNN_DEV const uint4* chunk_src(const ConvArgs& A, int c0, int y, int x) {
    if (MODE == 0) return (const uint4*)(A.a + ((size_t)y * A.W + x) * A.ca + c0);
    if (c0 < A.ca) return (const uint4*)(A.a + ((size_t)(y >> 1) * A.wa + (x >> 1)) * A.ca + c0);
    return (const uint4*)(A.b + ((size_t)y * A.W + x) * A.cb + (c0 - A.ca));
}

template <int NG, int NTW, int MT, int MODE, bool ONE_CHUNK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SR_PIPE_WAVES)))
k_conv3x3_pipe(ConvArgs A, int tilesX, int tiles) {
    static_assert(MODE == 0 || MODE == 2, "same-resolution or upsample + concatenation sources");
    constexpr int kRG = 4 / NG;
    static_assert(NG * kRG == 4, "4 waves split into channel groups x row groups");
    constexpr int kCo = NG * 16 * NTW;
    constexpr int kRows = kRG * MT;
    constexpr int kHaloPx = kHW * (kRows + 2);
    constexpr int kHaloWords = kHaloPx * 4;
    constexpr int kIter = (kHaloWords + 255) / 256;
    constexpr int kWWords = ONE_CHUNK ? 0 : 9 * kCo * 4;  // weight rows [tap][channel], 4 words each
    constexpr int kWIter = (kWWords + 255) / 256;
    __shared__ uint4 smem[kHaloWords + kWWords];
    uint4* const halo = smem;
    uint4* const wlds = smem + kHaloWords;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cg = wave % NG, rg = wave / NG;
    const int col = lane & 15, grp = lane >> 4;
    const int chunks = ONE_CHUNK ? 1 : A.cinPad >> 5;
    bf16x8 wf[9][NTW];
    if (ONE_CHUNK) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
#pragma unroll
            for (int h = 0; h < NTW; ++h)
                wf[tap][h] = __builtin_bit_cast(
                    bf16x8, *(const uint4*)(A.w + ((size_t)tap * A.coutPad + (cg * NTW + h) * 16 + col) * A.cinPad +
                                            grp * 8));
    }
    float bias[NTW][4];
#pragma unroll
    for (int h = 0; h < NTW; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = (cg * NTW + h) * 16 + grp * 4 + r;
            bias[h][r] = co < A.cout ? A.bias[co] : 0.0f;
        }
    uint4 st[kIter];
    uint4 sw[kWIter > 0 ? kWIter : 1];
    // Step (t, c) into registers. Halo word idx = threadIdx.x + 256 i is channel group idx & 3 of
    // halo pixel idx >> 2, zero outside the image; weight word e is group e & 3 of row e >> 2 =
    // tap * kCo + channel.
    auto load_step = [&](int t, int c) {
        const int tx0 = (t % tilesX) * kTW - 1, ty0 = (t / tilesX) * kRows - 1;
#pragma unroll
        for (int i = 0; i < kIter; ++i) {
            const int idx = threadIdx.x + 256 * i;
            const int p = idx >> 2, part = idx & 3;
            const int y = ty0 + p / kHW, x = tx0 + p % kHW;
            const bool ok = idx < kHaloWords && x >= 0 && y >= 0 && x < A.W && y < A.H;
            const uint4 v = chunk_src<MODE>(A, c * 32 + part * 8, ok ? y : 0, ok ? x : 0)[0];
            st[i] = ok ? v : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < kWIter; ++i) {
            const int e = min((int)threadIdx.x + 256 * i, kWWords - 1);
            const int row = e >> 2, tap = row / kCo, co = row % kCo;
            sw[i] = *(const uint4*)(A.w + ((size_t)tap * A.coutPad + co) * A.cinPad + c * 32 + (e & 3) * 8);
        }
    };
    int t = blockIdx.x, c = 0;
    if (t < tiles) load_step(t, 0);
    f32x4 acc[MT][NTW];
    while (t < tiles) {
        __syncthreads();  // the previous step's fragments are read
#pragma unroll
        for (int i = 0; i < kIter; ++i) {
            const int idx = threadIdx.x + 256 * i;
            if (idx < kHaloWords) halo[lds_slot(idx >> 2, idx & 3)] = st[i];
        }
#pragma unroll
        for (int i = 0; i < kWIter; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (e < kWWords) wlds[lds_slot(e >> 2, e & 3)] = sw[i];
        }
        __syncthreads();
        const int tn = c + 1 < chunks ? t : t + (int)gridDim.x, cn = c + 1 < chunks ? c + 1 : 0;
        if (tn < tiles) load_step(tn, cn);
        if (!ONE_CHUNK) {
#pragma unroll
            for (int tap = 0; tap < 9; ++tap)
#pragma unroll
                for (int h = 0; h < NTW; ++h)
                    wf[tap][h] = __builtin_bit_cast(bf16x8, wlds[lds_slot(tap * kCo + (cg * NTW + h) * 16 + col, grp)]);
        }
        if (c == 0) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int h = 0; h < NTW; ++h) acc[mt][h] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int R = 0; R < MT + 2; ++R)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int p = (rg * MT + R) * kHW + col + dx;
                const bf16x8 px = __builtin_bit_cast(bf16x8, halo[lds_slot(p, grp)]);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    const int mt = R - dy;
                    if (mt < 0 || mt >= MT) continue;
#pragma unroll
                    for (int h = 0; h < NTW; ++h)
                        acc[mt][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[dy * 3 + dx][h], px, acc[mt][h], 0, 0, 0);
                }
            }
        if (c + 1 == chunks) {
            // acc[mt][h][r]: channel (cg NTW + h) 16 + 4 grp + r of pixel (tx0 + col, ty0 + rg MT + mt).
            const int x = (t % tilesX) * kTW + col, ty0 = (t / tilesX) * kRows;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int y = ty0 + rg * MT + mt;
                if (x >= A.W || y >= A.H) continue;
#pragma unroll
                for (int h = 0; h < NTW; ++h) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int co = (cg * NTW + h) * 16 + grp * 4 + r;
                        v[r] = acc[mt][h][r] + bias[h][r];
                        if (A.relu) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
                        if (co >= A.cout) v[r] = 0.0f;
                    }
                    *(uint2*)out_at(A, (size_t)y * A.W + x, (cg * NTW + h) * 16 + grp * 4) =
                        make_uint2(pack2(v[0], v[1]), pack2(v[2], v[3]));
                }
            }
        }
        t = tn;
        c = cn;
    }
}

// Input features: log(1 + radiance), albedo, normal, 1 / (1 + depth) (0 without a surface).
__global__ void k_features(const float* rad, const float* alb, const float* nrm, const float* dep, int W, int H,
                           uint16_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    float f[16];
    for (int c = 0; c < 3; ++c) f[c] = log1pf(fmaxf(rad[i * 3 + c], 0.0f));
    for (int c = 0; c < 3; ++c) f[3 + c] = alb[i * 3 + c];
    for (int c = 0; c < 3; ++c) f[6 + c] = nrm[i * 3 + c];
    const float d = dep[i];
    f[9] = isinf(d) ? 0.0f : 1.0f / (1.0f + d);
    for (int c = kFeat; c < 16; ++c) f[c] = 0.0f;
    uint4* o = (uint4*)(out + (size_t)i * 32);
    o[0] = make_uint4(pack2(f[0], f[1]), pack2(f[2], f[3]), pack2(f[4], f[5]), pack2(f[6], f[7]));
    o[1] = make_uint4(pack2(f[8], f[9]), 0u, 0u, 0u);  // channels 16..31 stay zero from allocation
}

// One stage of the kernel cascade: out(p) = sum_k w_k(p) in(clamp(p + d (dx_k, dy_k))), w = softmax of
// the stage's 25 logits, 5 x 5 taps spaced d = 2^stage pixels, edges clamped.
// The logits are stage-planar ([stage][pixel][32 channels]); the stage's 25 sit at the start of the
// pixel's 64-byte row in plane STAGE and are read as four 16-byte words.
// A 256-thread workgroup owns a TS x TS pixel tile (TS * TS / 256 pixels a thread) and first
// stages the input over the tile and 2d pixels around it in LDS, one plane per colour channel,
// with the image-edge clamp applied as it loads; each pixel then reads its 25 taps from LDS.
template <int STAGE, int TS>
__global__ void __launch_bounds__(256) k_kpn_stage(const uint16_t* logits, const float* in, int W, int H,
                                                   float* out) {
    constexpr int kD = 1 << STAGE;
    constexpr int kSide = TS + 4 * kD;  // tile and 2d halo pixels either side
    constexpr int kPx = kSide * kSide;
    constexpr int kPPT = TS * TS / 256;
    static_assert(kPPT * 256 == TS * TS, "whole pixels per thread");
    __shared__ float plane[3][kPx];
    const int tx0 = blockIdx.x * TS, ty0 = blockIdx.y * TS;
    constexpr int kO = 0;
    static_assert(kTaps <= kStageStride && kStages * kStageStride == kLogitStride, "one 64-byte row a stage");
    const uint16_t* lplane = logits + (size_t)STAGE * W * H * kStageStride;
    // The thread's logits first (they do not wait for the staged input), then the staging loads in
    // batches of 8 ahead of their LDS stores.
    uint32_t wds[kPPT][16];
#pragma unroll
    for (int q = 0; q < kPPT; ++q) {
        const int lp = threadIdx.x + 256 * q;
        const int x = min(tx0 + lp % TS, W - 1), y = min(ty0 + lp / TS, H - 1);
        const uint4* src = (const uint4*)(lplane + ((size_t)y * W + x) * kStageStride);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint4 v = src[j];
            wds[q][4 * j] = v.x; wds[q][4 * j + 1] = v.y; wds[q][4 * j + 2] = v.z; wds[q][4 * j + 3] = v.w;
        }
    }
    // Row r of the staged area is image row clamp(ty0 - 2d + r); its 3 kSide floats are read as one
    // run of columns clamp(tx0 - 2d + c).
    constexpr int kF = kPx * 3, kFIt = (kF + 255) / 256, kB = 8;
#pragma unroll
    for (int i0 = 0; i0 < kFIt; i0 += kB) {
        float v[kB];
#pragma unroll
        for (int j = 0; j < kB; ++j) {
            if (i0 + j >= kFIt) break;
            const int f = min((int)threadIdx.x + 256 * (i0 + j), kF - 1);
            const int r = f / (kSide * 3), rem = f - r * (kSide * 3), cl = rem / 3, ch = rem - cl * 3;
            const int y = min(max(ty0 - 2 * kD + r, 0), H - 1), x = min(max(tx0 - 2 * kD + cl, 0), W - 1);
            v[j] = in[((size_t)y * W + x) * 3 + ch];
        }
#pragma unroll
        for (int j = 0; j < kB; ++j) {
            if (i0 + j >= kFIt) break;
            const int f = (int)threadIdx.x + 256 * (i0 + j);
            const int r = f / (kSide * 3), rem = f - r * (kSide * 3), cl = rem / 3, ch = rem - cl * 3;
            if (f < kF) plane[ch][r * kSide + cl] = v[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPPT; ++q) {
        const int lp = threadIdx.x + 256 * q;
        const int lx = lp % TS, ly = lp / TS;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= W || y >= H) continue;
        const int i = y * W + x;
        float l[kTaps];
        float m = -INFINITY;
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            const uint32_t wv = wds[q][(kO + k) >> 1];
            l[k] = bf2f((uint16_t)(((kO + k) & 1) ? (wv >> 16) : (wv & 0xffffu)));
            m = fmaxf(m, l[k]);
        }
        float s = 0.0f, r = 0.0f, g = 0.0f, b = 0.0f;
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            const float e = expf(l[k] - m);
            const int o = (ly + 2 * kD + kD * (k / 5 - 2)) * kSide + lx + 2 * kD + kD * (k % 5 - 2);
            s += e;
            r += e * plane[0][o];
            g += e * plane[1][o];
            b += e * plane[2][o];
        }
        out[i * 3] = r / s;
        out[i * 3 + 1] = g / s;
        out[i * 3 + 2] = b / s;
    }
}

// ---- Training kernels --------------------------------------------------------------------------------
// Loss mean (out - t)^2 / (t^2 + 0.01) and its gradient g = dL/dout.
__global__ void k_kpn_loss(const float* out, const float* target, int n, float invN, float* g, float* loss) {
    __shared__ float part[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.0f;
    if (i < n) {
        for (int c = 0; c < 3; ++c) {
            const float t = target[(size_t)i * 3 + c], d = out[(size_t)i * 3 + c] - t, den = t * t + 0.01f;
            lsum += d * d / den * invN;
            g[(size_t)i * 3 + c] = 2.0f * d / den * invN;
        }
    }
    part[threadIdx.x] = lsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// Backward of one cascade stage: with g = dL/dout(p), dL/dw_k = sum_c g_c in(q_k)_c and the softmax
// gives dL/dlogit_k = w_k (dL/dw_k - sum_j w_j dL/dw_j) (written as bf16 into the stage's logit
// channels); dL/din(q_k) += w_k g is scattered into gin when it is needed.
__global__ void k_kpn_stage_back(const uint16_t* logits, int stage, const float* in, const float* gout, int W, int H,
                                 uint16_t* dlog, float* gin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int x = i % W, y = i / W, d = 1 << stage;
    const uint16_t* lg = logits + ((size_t)stage * W * H + i) * kStageStride;  // stage-planar logits
    float w[kTaps];
    int q[kTaps];
    float m = -INFINITY;
    for (int k = 0; k < kTaps; ++k) {
        w[k] = bf2f(lg[k]);
        m = fmaxf(m, w[k]);
    }
    float s = 0.0f;
    for (int k = 0; k < kTaps; ++k) {
        w[k] = expf(w[k] - m);
        s += w[k];
    }
    const float g0 = gout[(size_t)i * 3], g1 = gout[(size_t)i * 3 + 1], g2 = gout[(size_t)i * 3 + 2];
    float dw[kTaps], dot = 0.0f;
    for (int k = 0; k < kTaps; ++k) {
        w[k] /= s;
        const int yy = min(max(y + d * (k / 5 - 2), 0), H - 1), xx = min(max(x + d * (k % 5 - 2), 0), W - 1);
        q[k] = yy * W + xx;
        const float* p = in + (size_t)q[k] * 3;
        dw[k] = g0 * p[0] + g1 * p[1] + g2 * p[2];
        dot += w[k] * dw[k];
    }
    uint16_t* dl = dlog + (size_t)i * kLogitStride + stage * kStageStride;
    for (int k = 0; k < kTaps; ++k) dl[k] = f2bf(w[k] * (dw[k] - dot));
    if (gin)
        for (int k = 0; k < kTaps; ++k) {
            float* t = gin + (size_t)q[k] * 3;
            unsafeAtomicAdd(&t[0], w[k] * g0);
            unsafeAtomicAdd(&t[1], w[k] * g1);
            unsafeAtomicAdd(&t[2], w[k] * g2);
        }
}

// Output gradient of a layer through its ReLU (kept where the bf16 activation is positive), as
// bf16; the fp32 accumulator is cleared for the next step.
__global__ void k_relu_back(float* dact, const uint16_t* act, size_t n, int relu, uint16_t* dy) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = dact[i];
    if (relu && !(bf2f(act[i]) > 0.0f)) g = 0.0f;
    dy[i] = f2bf(g);
    dact[i] = 0.0f;
}

// Gradient of a conv's staged input (bf16, stride cinPad, at the conv's resolution W x H) back to
// its sources. Mode 0: same pixels. Mode 1 (2x2 average of a finer source wa x ha, last row /
// column repeated): fine pixel (Y, X) receives 0.25 of its coarse pixel, twice along an axis
// where it was read twice. Mode 2: the first caPad channels go to the coarse source (sum over the
// 2x2 fine pixels that read it), the rest to the skip source.
__global__ void k_stage_back_a(const uint16_t* dx, int W, int H, int cinPad, int mode, float* da, int wa, int ha,
                               int caPad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)wa * ha * caPad) return;
    const int c = (int)(i % caPad);
    const int p = (int)(i / caPad);
    const int X = p % wa, Y = p / wa;
    float g = 0.0f;
    if (mode == 0) {
        g = bf2f(dx[(size_t)p * cinPad + c]);
    } else if (mode == 1) {
        const float mx = (X == wa - 1 && (wa & 1)) ? 2.0f : 1.0f, my = (Y == ha - 1 && (ha & 1)) ? 2.0f : 1.0f;
        g = 0.25f * mx * my * bf2f(dx[((size_t)(Y >> 1) * W + (X >> 1)) * cinPad + c]);
    } else {
        for (int dy = 0; dy < 2; ++dy)
            for (int dxx = 0; dxx < 2; ++dxx) {
                const int yy = 2 * Y + dy, xx = 2 * X + dxx;
                if (yy < H && xx < W) g += bf2f(dx[((size_t)yy * W + xx) * cinPad + c]);
            }
    }
    da[i] += g;
}
__global__ void k_stage_back_b(const uint16_t* dx, int W, int H, int cinPad, int caPad, float* db, int cbPad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)W * H * cbPad) return;
    const int c = (int)(i % cbPad);
    const size_t p = i / cbPad;
    db[i] += bf2f(dx[p * cinPad + caPad + c]);
}

// Weight gradient of one layer: dW[tap][co][ci] = sum over pixels of dY[p][co] X[p + off(tap)][ci],
// X being the conv's staged input. A workgroup owns one (tap, 32-channel chunk) pair and a strided
// set of 16 x 8 pixel tiles; it stages dY (transposed, [co][pixel]) and the input halo
// (transposed, [ci][halo pixel]) in LDS, and each wave multiplies its 32 pixels of the tile:
// A = dY^T (16 co x 32 px), B = X (32 px x 16 ci) on v_mfma_f32_16x16x32_bf16.
template <int MT>
__global__ void __launch_bounds__(256) k_conv_wgrad(ConvArgs X, const uint16_t* dy, float* gw) {
    __shared__ uint16_t dyT[MT * 16][kTW * kTH];
    __shared__ uint16_t xT[32][kHalo];
    const int tap = blockIdx.x % 9, c0 = (blockIdx.x / 9) * 32;
    const int ky = tap / 3, kx = tap % 3;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int coutPad = X.coutPad;  // channel stride of dy
    f32x4 acc[MT][2];
    for (int mt = 0; mt < MT; ++mt)
        for (int nt = 0; nt < 2; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int tilesX = (X.W + kTW - 1) / kTW, tilesY = (X.H + kTH - 1) / kTH;
    for (int t = blockIdx.y; t < tilesX * tilesY; t += gridDim.y) {
        const int tx0 = (t % tilesX) * kTW, ty0 = (t / tilesX) * kTH;
        __syncthreads();
        for (int idx = threadIdx.x; idx < kTW * kTH * MT * 2; idx += 256) {
            const int px = idx / (MT * 2), g = idx % (MT * 2);
            const int y = ty0 + px / kTW, x = tx0 + px % kTW;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (x < X.W && y < X.H) v = *(const uint4*)(dy + ((size_t)y * X.W + x) * coutPad + g * 8);
            const uint32_t q[4] = {v.x, v.y, v.z, v.w};
            for (int j = 0; j < 8; ++j) dyT[g * 8 + j][px] = (uint16_t)(q[j >> 1] >> ((j & 1) * 16));
        }
        for (int idx = threadIdx.x; idx < kHalo * 4; idx += 256) {
            const int p = idx >> 2, part = idx & 3;
            const uint4 v = fetch8(X, c0 + part * 8, ty0 - 1 + p / kHW, tx0 - 1 + p % kHW);
            const uint32_t q[4] = {v.x, v.y, v.z, v.w};
            for (int j = 0; j < 8; ++j) xT[part * 8 + j][p] = (uint16_t)(q[j >> 1] >> ((j & 1) * 16));
        }
        __syncthreads();
        const int k0 = wave * 32 + (lane >> 4) * 8;  // this lane's 8 pixels (one tile row)
        const int h0 = (k0 / kTW + ky) * kHW + k0 % kTW + kx;
        bf16x8 bfr[2];
        for (int nt = 0; nt < 2; ++nt) {
            const uint16_t* src = &xT[nt * 16 + (lane & 15)][h0];
            uint32_t q[4];
            for (int j = 0; j < 4; ++j) q[j] = (uint32_t)src[2 * j] | ((uint32_t)src[2 * j + 1] << 16);
            bfr[nt] = __builtin_bit_cast(bf16x8, make_uint4(q[0], q[1], q[2], q[3]));
        }
        for (int mt = 0; mt < MT; ++mt) {
            const bf16x8 afr = __builtin_bit_cast(bf16x8, *(const uint4*)&dyT[mt * 16 + (lane & 15)][k0]);
            for (int nt = 0; nt < 2; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nt], acc[mt][nt], 0, 0, 0);
        }
    }
    for (int mt = 0; mt < MT; ++mt)
        for (int nt = 0; nt < 2; ++nt)
            for (int r = 0; r < 4; ++r) {
                const float v = acc[mt][nt][r];
                if (v == 0.0f) continue;
                const int co = mt * 16 + 4 * (lane >> 4) + r, ci = c0 + nt * 16 + (lane & 15);
                unsafeAtomicAdd(&gw[((size_t)tap * X.coutPad + co) * X.cinPad + ci], v);
            }
}

// Bias gradient: sum of dY over pixels, one channel per lane, 256 pixels per block.
__global__ void k_bias_grad(const uint16_t* dy, int n, int coutPad, float* gb) {
    const int c = threadIdx.x;
    if (c >= coutPad) return;
    const int p0 = blockIdx.x * 256, p1 = min(p0 + 256, n);
    float s = 0.0f;
    for (int p = p0; p < p1; ++p) s += bf2f(dy[(size_t)p * coutPad + c]);
    if (s != 0.0f) unsafeAtomicAdd(&gb[c], s);
}

// Adam (bias-corrected), gradients cleared after use.
__global__ void k_adam(float* w, float* g, float* m1, float* m2, size_t n, float lr, float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi, b = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
    g[i] = 0.0f;
}

// bf16 forward weights [tap][co][ci] and flipped-transposed data-gradient weights
// [8 - tap][ci][co] from the fp32 masters.
__global__ void k_pack_weights(const float* mw, int coutPad, int cinPad, uint16_t* w, uint16_t* wt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)9 * coutPad * cinPad) return;
    const int ci = (int)(i % cinPad);
    const int co = (int)((i / cinPad) % coutPad);
    const int tap = (int)(i / ((size_t)cinPad * coutPad));
    const uint16_t b = f2bf(mw[i]);
    w[i] = b;
    wt[((size_t)(8 - tap) * cinPad + ci) * coutPad + co] = b;
}

uint16_t host_f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float host_bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

void level_dims(uint32_t W, uint32_t H, int level, uint32_t& w, uint32_t& h) {
    w = W;
    h = H;
    for (int l = 0; l < level; ++l) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
}

int cin_pad(const CnnLayerSpec& L, int srcAPad, int srcBPad) {
    return L.mode == 2 ? srcAPad + srcBPad : srcAPad;
}

}  // namespace

const std::vector<CnnLayerSpec>& cnn_layers() {
    static const std::vector<CnnLayerSpec> L = {
        {kFeat, 0, 32, 0, -1, -1, 0, true},   // 0
        {32, 0, 32, 0, 0, -1, 0, true},       // 1  skip 0
        {32, 0, 48, 1, 1, -1, 1, true},       // 2
        {48, 0, 48, 0, 2, -1, 1, true},       // 3  skip 1
        {48, 0, 64, 1, 3, -1, 2, true},       // 4
        {64, 0, 64, 0, 4, -1, 2, true},       // 5  skip 2
        {64, 0, 96, 1, 5, -1, 3, true},       // 6
        {96, 0, 96, 0, 6, -1, 3, true},       // 7  skip 3
        {96, 0, 128, 1, 7, -1, 4, true},      // 8
        {128, 0, 128, 0, 8, -1, 4, true},     // 9
        {128, 96, 96, 2, 9, 7, 3, true},      // 10
        {96, 0, 96, 0, 10, -1, 3, true},      // 11
        {96, 64, 64, 2, 11, 5, 2, true},      // 12
        {64, 0, 64, 0, 12, -1, 2, true},      // 13
        {64, 48, 48, 2, 13, 3, 1, true},      // 14
        {48, 0, 48, 0, 14, -1, 1, true},      // 15
        {48, 32, 32, 2, 15, 1, 0, true},      // 16
        {32, 0, 32, 0, 16, -1, 0, true},      // 17
        {32, 0, kLogits, 0, 17, -1, 0, false},  // 18 kernel logits of the four cascade stages
    };
    return L;
}

// Device channel of logical output channel co of layer li: the head's logit (stage s, tap k),
// logical channel 25 s + k, is channel 32 s + k; every other layer keeps its order.
static int packed_channel(int li, int co) {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    if (li != (int)Ls.size() - 1) return co;
    return kStageStride * (co / kTaps) + co % kTaps;
}

uint64_t cnn_weight_count() {
    uint64_t n = 0;
    for (const CnnLayerSpec& L : cnn_layers()) n += (uint64_t)L.cout * (L.cinA + L.cinB) * 9 + L.cout;
    return n;
}

double CnnNet::flops(uint32_t W, uint32_t H) {
    double f = 0.0;
    for (const CnnLayerSpec& L : cnn_layers()) {
        uint32_t w, h;
        level_dims(W, H, L.level, w, h);
        f += 2.0 * 9.0 * (double)(L.cinA + L.cinB) * L.cout * w * h;
    }
    return f;
}

CnnNet::~CnnNet() {
    for (uint16_t* p : w_) (void)hipFree(p);
    for (float* p : b_) (void)hipFree(p);
    release_acts();
    for (auto* v : {&mw_, &mb_, &gw_, &gb_, &m1w_, &m2w_, &m1b_, &m2b_})
        for (float* p : *v) (void)hipFree(p);
    for (uint16_t* p : wt_) (void)hipFree(p);
    if (zeros_) (void)hipFree(zeros_);
    release_train_acts();
}

void CnnNet::release_acts() {
    if (feat_) (void)hipFree(feat_);
    feat_ = nullptr;
    for (float*& p : casc_) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    for (uint16_t* p : act_) (void)hipFree(p);
    act_.clear();
    W_ = H_ = 0;
}

bool CnnNet::load(const float* weights, hipStream_t s) {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    for (uint16_t* p : w_) (void)hipFree(p);
    for (float* p : b_) (void)hipFree(p);
    w_.assign(Ls.size(), nullptr);
    b_.assign(Ls.size(), nullptr);
    loaded_ = false;
    const float* src = weights;
    for (size_t li = 0; li < Ls.size(); ++li) {
        const CnnLayerSpec& L = Ls[li];
        const int cin = L.cinA + L.cinB;
        const int padA = pad32(L.cinA), cinP = L.mode == 2 ? padA + pad32(L.cinB) : padA;
        const int coutP = pad32(L.cout);
        std::vector<uint16_t> w((size_t)9 * coutP * cinP, 0);
        for (int co = 0; co < L.cout; ++co)
            for (int ci = 0; ci < cin; ++ci) {
                const int cp = ci < L.cinA ? ci : padA + (ci - L.cinA);
                const int pc = packed_channel((int)li, co);
                for (int tap = 0; tap < 9; ++tap)
                    w[((size_t)tap * coutP + pc) * cinP + cp] = host_f2bf(src[((size_t)co * cin + ci) * 9 + tap]);
            }
        src += (size_t)L.cout * cin * 9;
        std::vector<float> b(coutP, 0.0f);
        for (int co = 0; co < L.cout; ++co) b[packed_channel((int)li, co)] = src[co];
        src += L.cout;
        if (hipMalloc((void**)&w_[li], w.size() * 2) != hipSuccess) return false;
        if (hipMalloc((void**)&b_[li], b.size() * 4) != hipSuccess) return false;
        if (hipMemcpyAsync(w_[li], w.data(), w.size() * 2, hipMemcpyHostToDevice, s) != hipSuccess) return false;
        if (hipMemcpyAsync(b_[li], b.data(), b.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) return false;
        if (hipStreamSynchronize(s) != hipSuccess) return false;
    }
    loaded_ = true;
    return true;
}

bool CnnNet::ensure(uint32_t W, uint32_t H) {
    if (W == W_ && H == H_ && feat_) return true;
    release_acts();
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    if (hipMalloc((void**)&feat_, (size_t)W * H * 32 * 2) != hipSuccess) return false;
    if (hipMemset(feat_, 0, (size_t)W * H * 32 * 2) != hipSuccess) return false;
    for (float*& p : casc_)
        if (hipMalloc((void**)&p, (size_t)W * H * 12) != hipSuccess) return false;
    act_.assign(Ls.size(), nullptr);
    for (size_t li = 0; li < Ls.size(); ++li) {
        uint32_t w, h;
        level_dims(W, H, Ls[li].level, w, h);
        const size_t bytes = (size_t)w * h * pad32(Ls[li].cout) * 2;
        if (hipMalloc((void**)&act_[li], bytes) != hipSuccess) return false;
        // Channels beyond cout are never written: they stay zero for the next layer's K chunks.
        if (hipMemset(act_[li], 0, bytes) != hipSuccess) return false;
    }
    W_ = W;
    H_ = H;
    return true;
}

namespace {
// Forward arguments of layer li at image size W x H.
ConvArgs layer_args(int li, uint32_t W, uint32_t H, const uint16_t* feat, const std::vector<uint16_t*>& act,
                    const std::vector<uint16_t*>& w, const std::vector<float*>& b) {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    const CnnLayerSpec& L = Ls[li];
    ConvArgs A;
    memset(&A, 0, sizeof(A));
    uint32_t wd, h;
    level_dims(W, H, L.level, wd, h);
    const int srcLevel = L.mode == 1 ? L.level - 1 : (L.mode == 2 ? L.level + 1 : L.level);
    uint32_t wa, ha;
    level_dims(W, H, srcLevel, wa, ha);
    A.a = L.srcA < 0 ? feat : act[L.srcA];
    A.ca = L.srcA < 0 ? 32 : pad32(Ls[L.srcA].cout);
    A.b = L.srcB < 0 ? nullptr : act[L.srcB];
    A.cb = L.srcB < 0 ? 0 : pad32(Ls[L.srcB].cout);
    A.wa = (int)wa;
    A.ha = (int)ha;
    A.mode = L.mode;
    A.W = (int)wd;
    A.H = (int)h;
    A.cinPad = cin_pad(L, A.ca, A.cb);
    A.coutPad = pad32(L.cout);
    // The head's logit slots are spread over all 128 channels; its padding channels have zero
    // weights and biases and come out zero.
    A.cout = li == (int)Ls.size() - 1 ? A.coutPad : L.cout;
    // The head writes its logits stage-planar, [stage][pixel][32], so that each cascade stage reads
    // whole 64-byte rows of one plane.
    A.plane = li == (int)Ls.size() - 1 ? (size_t)wd * h * kStageStride : 0;
    A.w = w.empty() ? nullptr : w[li];
    A.bias = b.empty() ? nullptr : b[li];
    A.out = act[li];
    A.relu = L.relu ? 1 : 0;
    return A;
}

#ifndef SR_CONV_SMALL_MT
#define SR_CONV_SMALL_MT 4
#endif
constexpr int kSmallMT = SR_CONV_SMALL_MT;  // rows per wave of the 32-channel (2-tile) layers
// Rows per wave by output-channel tiles, measured per layer at 1080p (more rows reuse each weight
// fragment more often, more tiles hold more accumulators): 2 tiles 4 rows (8 rows: 94 -> 113 us on
// L0), 3 tiles 8 rows (4 rows: L14 143 vs 104 us), 4 tiles 4 rows (2 rows: L5 43 vs 30 us), 6 and 8
// tiles 2 rows (4 rows: L10 52 vs 64 us), the 7-tile head 4 rows (2 rows: 410 vs 345 us).
#ifndef SR_HEAD_MT
#define SR_HEAD_MT 4
#endif
// Layers of 64 to 128 output channels run as k_conv3x3_split (each wave owns 32 output channels
// of the block's tile), measured per layer at 1080p against k_conv3x3: the 100-logit head with 4
// rows 295 vs 349 us (8 rows 318 us), 128 channels with 4 rows L9 22 vs 45 us, 96 channels with 4
// rows L10 42 vs 53 us (8 rows 48 us), 64 channels with 8 rows L5 22 vs 30 us (4 rows 28 us). The
// 48-channel layers stay on k_conv3x3 (split into 2 waves with 16 padded channels: L14 140 vs 105 us).
#ifndef SR_HEAD_SPLIT
#define SR_HEAD_SPLIT 1
#endif
#ifndef SR_HEAD_SPLIT_MT
#define SR_HEAD_SPLIT_MT 4
#endif
#ifndef SR_SPLIT128
#define SR_SPLIT128 1
#endif
#ifndef SR_SPLIT96
#define SR_SPLIT96 1
#endif
#ifndef SR_SPLIT64
#define SR_SPLIT64 1
#endif
#ifndef SR_SPLIT96_MT
#define SR_SPLIT96_MT 4
#endif
#ifndef SR_SPLIT64_MT
#define SR_SPLIT64_MT 8
#endif
#ifndef SR_MT3
#define SR_MT3 8
#endif
#ifndef SR_MT4
#define SR_MT4 4
#endif
#ifndef SR_MT6
#define SR_MT6 2
#endif
#ifndef SR_MT8
#define SR_MT8 2
#endif

template <int MT, int NW>
void launch_split(const ConvArgs& A, hipStream_t s) {
    const dim3 grid((A.W + kTW - 1) / kTW, (A.H + MT - 1) / MT);
    hipLaunchKernelGGL((k_conv3x3_split<MT, NW>), grid, dim3(NW * 64), 0, s, A);
}

template <int NT, int MT>
void launch_conv_t(const ConvArgs& A, hipStream_t s) {
    const dim3 grid((A.W + kTW - 1) / kTW, (A.H + 4 * MT - 1) / (4 * MT));
    hipLaunchKernelGGL((k_conv3x3<NT, MT>), grid, dim3(256), 0, s, A);
}

// Same-resolution layers of one 32-channel K chunk (the full-resolution 32-channel layers and the
// head) run as k_conv3x3_pipe, measured at 1080p against k_conv3x3 / k_conv3x3_split: 32 channels
// 55-60 us vs 95-108 us, the head 197 us vs 300 us. Rows per wave and head channel groups measured:
// 2 rows 59-65 us on the 32-channel layers; head with 2 rows 254 us, with 2 channel groups of 64
// (4 or 2 rows) 250 / 286 us, against 238 us for 4 groups and 4 rows (before stage-planar logits).
#ifndef SR_RES
#define SR_RES 1
#endif
#ifndef SR_RES32_MT
#define SR_RES32_MT 4
#endif
#ifndef SR_RES_HEAD_NG
#define SR_RES_HEAD_NG 4
#endif
#ifndef SR_RES_HEAD_MT
#define SR_RES_HEAD_MT 4
#endif
// The upsampling 32-channel layer (3 K chunks) through k_conv3x3_pipe, measured at 1080p: 272 us
// with the chunk weights staged in LDS, 223 us with them loaded into registers per step, against
// 209 us on k_conv3x3<2, 4> with batched halo loads (3 waves per SIMD).
#ifndef SR_PIPE_UP
#define SR_PIPE_UP 0
#endif

#ifndef SR_PIPE_UP_MT
#define SR_PIPE_UP_MT 4
#endif

// One resident workgroup per slot the occupancy calculator reports on every CU, at most one per tile.
template <int NG, int NTW, int MT, int MODE, bool ONE_CHUNK>
void launch_pipe(const ConvArgs& A, hipStream_t s) {
    constexpr int kRows = (4 / NG) * MT;
    const int tilesX = (A.W + kTW - 1) / kTW, tiles = tilesX * ((A.H + kRows - 1) / kRows);
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_conv3x3_pipe<NG, NTW, MT, MODE, ONE_CHUNK>, 256,
                                                       0);
    const int grid = std::max(1, std::min(tiles, std::max(1, cus) * std::max(1, per)));
    hipLaunchKernelGGL((k_conv3x3_pipe<NG, NTW, MT, MODE, ONE_CHUNK>), dim3(grid), dim3(256), 0, s, A, tilesX, tiles);
}

// Few output channels: 8 rows per wave (weights reused 8 times, 16-byte stores); more: 2 rows.
bool launch_conv(const ConvArgs& A, hipStream_t s) {
    if (SR_RES && A.mode == 0 && A.cinPad == 32 && A.ca == 32) {
        if (A.coutPad == 32) {
            launch_pipe<1, 2, SR_RES32_MT, 0, true>(A, s);
            return true;
        }
        if (A.coutPad == 128) {
            launch_pipe<SR_RES_HEAD_NG, 128 / 16 / SR_RES_HEAD_NG, SR_RES_HEAD_MT, 0, true>(A, s);
            return true;
        }
    }
    if (SR_PIPE_UP && A.mode == 2 && A.coutPad == 32 && A.cout == 32) {
        launch_pipe<1, 2, SR_PIPE_UP_MT, 2, false>(A, s);
        return true;
    }
    switch ((A.cout + 15) / 16) {
        case 2: launch_conv_t<2, kSmallMT>(A, s); break;
        case 3: launch_conv_t<3, SR_MT3>(A, s); break;
        case 4:
            if (SR_SPLIT64 && A.coutPad == 64) launch_split<SR_SPLIT64_MT, 2>(A, s);
            else launch_conv_t<4, SR_MT4>(A, s);
            break;
        case 5: launch_conv_t<5, 2>(A, s); break;
        case 6:
            if (SR_SPLIT96 && A.coutPad == 96) launch_split<SR_SPLIT96_MT, 3>(A, s);
            else launch_conv_t<6, SR_MT6>(A, s);
            break;
        case 7:
            if (SR_HEAD_SPLIT && A.coutPad == 128) launch_split<SR_HEAD_SPLIT_MT, 4>(A, s);
            else launch_conv_t<7, SR_HEAD_MT>(A, s);
            break;
        case 8:
            if (SR_SPLIT128 && A.coutPad == 128) launch_split<SR_HEAD_SPLIT_MT, 4>(A, s);
            else launch_conv_t<8, SR_MT8>(A, s);
            break;
        case 10: launch_conv_t<10, 2>(A, s); break;
        case 14: launch_conv_t<14, 2>(A, s); break;
        default: return false;
    }
    return true;
}
}  // namespace

bool CnnNet::run(const float* rad, const float* alb, const float* nrm, const float* dep, uint32_t W, uint32_t H,
                 float* out, hipStream_t s) {
    if (!loaded_ || !ensure(W, H)) return false;
    const int n = (int)(W * H);
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, s, rad, alb, nrm, dep, (int)W, (int)H, feat_);
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    for (size_t li = 0; li < Ls.size(); ++li)
        if (!launch_conv(layer_args((int)li, W, H, feat_, act_, w_, b_), s)) return false;
    if (out) {
        static_assert(kStages == 4, "one k_kpn_stage instance per stage");
        void (*const stage[kStages])(const uint16_t*, const float*, int, int, float*) = {
            k_kpn_stage<0, 16>, k_kpn_stage<1, 16>, k_kpn_stage<2, 32>, k_kpn_stage<3, 32>};
        const int side[kStages] = {16, 16, 32, 32};
        for (int st = 0; st < kStages; ++st) {
            const dim3 g((W + side[st] - 1) / side[st], (H + side[st] - 1) / side[st]);
            hipLaunchKernelGGL(stage[st], g, dim3(256), 0, s, act_.back(), st ? casc_[st - 1] : rad, (int)W, (int)H,
                               st == kStages - 1 ? out : casc_[st]);
        }
    }
    return hipGetLastError() == hipSuccess;
}

bool CnnNet::read_layer(int layer, std::vector<float>& out, uint32_t& c, uint32_t& w, uint32_t& h) const {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    if (layer >= kGradLayerBase) {  // masked output gradient of a layer from the last training step
        const int li = layer - kGradLayerBase;
        if (dyp_.empty() || li >= (int)Ls.size()) return false;
        level_dims(tW_, tH_, Ls[li].level, w, h);
        c = Ls[li].cout;
        const int stride = pad32(Ls[li].cout);
        std::vector<uint16_t> raw((size_t)w * h * stride);
        if (hipMemcpy(raw.data(), dyp_[li], raw.size() * 2, hipMemcpyDeviceToHost) != hipSuccess) return false;
        out.resize((size_t)w * h * c);
        for (size_t p = 0; p < (size_t)w * h; ++p)
            for (uint32_t k = 0; k < c; ++k) out[p * c + k] = host_bf2f(raw[p * stride + packed_channel(li, (int)k)]);
        return true;
    }
    if (!feat_ || layer < -1 || layer >= (int)Ls.size()) return false;
    const int level = layer < 0 ? 0 : Ls[layer].level;
    level_dims(W_, H_, level, w, h);
    c = layer < 0 ? kFeat : Ls[layer].cout;
    const int stride = layer < 0 ? 32 : pad32(Ls[layer].cout);
    std::vector<uint16_t> raw((size_t)w * h * stride);
    if (hipMemcpy(raw.data(), layer < 0 ? feat_ : act_[layer], raw.size() * 2, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    out.resize((size_t)w * h * c);
    const bool planar = layer == (int)Ls.size() - 1;  // the head's stage-planar logits
    for (size_t p = 0; p < (size_t)w * h; ++p)
        for (uint32_t k = 0; k < c; ++k) {
            const int pc = layer < 0 ? (int)k : packed_channel(layer, (int)k);
            const size_t at = planar ? (size_t)(pc >> 5) * w * h * 32 + p * 32 + (pc & 31) : p * stride + pc;
            out[p * c + k] = host_bf2f(raw[at]);
        }
    return true;
}


// ---- Training ----------------------------------------------------------------------------------------
namespace {
size_t layer_weights(int li) {
    const CnnLayerSpec& L = cnn_layers()[li];
    const int padA = pad32(L.cinA), cinP = L.mode == 2 ? padA + pad32(L.cinB) : padA;
    return (size_t)9 * pad32(L.cout) * cinP;
}
int layer_cin_pad(int li) {
    const CnnLayerSpec& L = cnn_layers()[li];
    const int padA = pad32(L.cinA);
    return L.mode == 2 ? padA + pad32(L.cinB) : padA;
}
template <class T>
bool dalloc(T*& p, size_t n) {
    if (hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    return hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess;
}
}  // namespace

bool CnnNet::train_init(hipStream_t s) {
    if (!loaded_) return false;
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    const size_t nl = Ls.size();
    mw_.assign(nl, nullptr); mb_.assign(nl, nullptr); gw_.assign(nl, nullptr); gb_.assign(nl, nullptr);
    m1w_.assign(nl, nullptr); m2w_.assign(nl, nullptr); m1b_.assign(nl, nullptr); m2b_.assign(nl, nullptr);
    wt_.assign(nl, nullptr);
    for (size_t li = 0; li < nl; ++li) {
        const size_t nw = layer_weights((int)li), nb = (size_t)pad32(Ls[li].cout);
        if (!dalloc(mw_[li], nw) || !dalloc(gw_[li], nw) || !dalloc(m1w_[li], nw) || !dalloc(m2w_[li], nw) ||
            !dalloc(mb_[li], nb) || !dalloc(gb_[li], nb) || !dalloc(m1b_[li], nb) || !dalloc(m2b_[li], nb) ||
            !dalloc(wt_[li], nw))
            return false;
        // Masters from the loaded bf16 weights (exact) and fp32 biases.
        std::vector<uint16_t> wb(nw);
        std::vector<float> wf(nw);
        if (hipMemcpy(wb.data(), w_[li], nw * 2, hipMemcpyDeviceToHost) != hipSuccess) return false;
        for (size_t i = 0; i < nw; ++i) wf[i] = host_bf2f(wb[i]);
        if (hipMemcpy(mw_[li], wf.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess) return false;
        if (hipMemcpy(mb_[li], b_[li], nb * 4, hipMemcpyDeviceToDevice) != hipSuccess) return false;
    }
    if (!dalloc(zeros_, 256)) return false;
    step_ = 0;
    return pack(s);
}

bool CnnNet::pack(hipStream_t s) {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    for (size_t li = 0; li < Ls.size(); ++li) {
        const size_t nw = layer_weights((int)li);
        hipLaunchKernelGGL(k_pack_weights, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, mw_[li],
                           pad32(Ls[li].cout), layer_cin_pad((int)li), w_[li], wt_[li]);
    }
    return hipGetLastError() == hipSuccess;
}

void CnnNet::release_train_acts() {
    for (float* p : dact_) (void)hipFree(p);
    for (uint16_t* p : dyp_) (void)hipFree(p);
    dact_.clear();
    dyp_.clear();
    if (dx_) (void)hipFree(dx_);
    if (out_) (void)hipFree(out_);
    dx_ = nullptr;
    out_ = nullptr;
    for (float*& p : gcas_) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    tW_ = tH_ = 0;
}

bool CnnNet::ensure_train(uint32_t W, uint32_t H) {
    if (W == tW_ && H == tH_ && !dact_.empty()) return true;
    release_train_acts();
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    dact_.assign(Ls.size(), nullptr);
    dyp_.assign(Ls.size(), nullptr);
    size_t dxMax = 0;
    for (size_t li = 0; li < Ls.size(); ++li) {
        uint32_t w, h;
        level_dims(W, H, Ls[li].level, w, h);
        const size_t n = (size_t)w * h * pad32(Ls[li].cout);
        if (!dalloc(dact_[li], n) || !dalloc(dyp_[li], n)) return false;
        dxMax = std::max(dxMax, (size_t)w * h * layer_cin_pad((int)li));
    }
    if (!dalloc(dx_, dxMax) || !dalloc(out_, (size_t)W * H * 3)) return false;
    for (float*& p : gcas_)
        if (!dalloc(p, (size_t)W * H * 3)) return false;
    tW_ = W;
    tH_ = H;
    return true;
}

bool CnnNet::train_step(const float* rad, const float* alb, const float* nrm, const float* dep, const float* target,
                        uint32_t W, uint32_t H, float lr, float* lossDev, hipStream_t s, bool apply) {
    if (!training() || !ensure_train(W, H) || !run(rad, alb, nrm, dep, W, H, out_, s)) return false;
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    const int nl = (int)Ls.size();
    const int n = (int)(W * H);
    const dim3 g1((n + 255) / 256);
    hipLaunchKernelGGL(k_kpn_loss, g1, dim3(256), 0, s, out_, target, n, 1.0f / (3.0f * (float)n),
                       gcas_[kStages - 1], lossDev);
    for (int st = 0; st < kStages - 1; ++st)
        if (hipMemsetAsync(gcas_[st], 0, (size_t)n * 12, s) != hipSuccess) return false;
    for (int st = kStages - 1; st >= 0; --st)
        hipLaunchKernelGGL(k_kpn_stage_back, g1, dim3(256), 0, s, act_[nl - 1], st, st ? casc_[st - 1] : rad,
                           gcas_[st], (int)W, (int)H, dyp_[nl - 1], st ? gcas_[st - 1] : (float*)nullptr);
    for (int li = nl - 1; li >= 0; --li) {
        const CnnLayerSpec& L = Ls[li];
        const ConvArgs F = layer_args(li, W, H, feat_, act_, w_, b_);
        const size_t nact = (size_t)F.W * F.H * F.coutPad;
        if (li < nl - 1)
            hipLaunchKernelGGL(k_relu_back, dim3((unsigned)((nact + 255) / 256)), dim3(256), 0, s, dact_[li], act_[li],
                               nact, L.relu ? 1 : 0, dyp_[li]);
        // Weight and bias gradients.
        const dim3 wg(9 * (F.cinPad / 32), std::min(64, ((F.W + kTW - 1) / kTW) * ((F.H + kTH - 1) / kTH)));
        switch (F.coutPad / 16) {
            case 2: hipLaunchKernelGGL(k_conv_wgrad<2>, wg, dim3(256), 0, s, F, dyp_[li], gw_[li]); break;
            case 4: hipLaunchKernelGGL(k_conv_wgrad<4>, wg, dim3(256), 0, s, F, dyp_[li], gw_[li]); break;
            case 6: hipLaunchKernelGGL(k_conv_wgrad<6>, wg, dim3(256), 0, s, F, dyp_[li], gw_[li]); break;
            case 8: hipLaunchKernelGGL(k_conv_wgrad<8>, wg, dim3(256), 0, s, F, dyp_[li], gw_[li]); break;
            default: return false;
        }
        const int np = F.W * F.H;
        hipLaunchKernelGGL(k_bias_grad, dim3((np + 255) / 256), dim3(256), 0, s, dyp_[li], np, F.coutPad, gb_[li]);
        if (li == 0) break;  // no gradient for the input features
        // Data gradient: the same conv with flipped, transposed weights, into the staged-input space.
        ConvArgs D;
        memset(&D, 0, sizeof(D));
        D.a = dyp_[li];
        D.ca = F.coutPad;
        D.wa = F.W;
        D.ha = F.H;
        D.mode = 0;
        D.W = F.W;
        D.H = F.H;
        D.cinPad = F.coutPad;
        D.coutPad = F.cinPad;
        D.cout = F.cinPad;
        D.w = wt_[li];
        D.bias = zeros_;
        D.out = dx_;
        D.relu = 0;
        if (!launch_conv(D, s)) return false;
        // Back through pooling / upsampling / concatenation into the sources' gradients.
        if (L.srcA >= 0) {
            const size_t na = (size_t)F.wa * F.ha * F.ca;
            hipLaunchKernelGGL(k_stage_back_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, dx_, F.W, F.H,
                               F.cinPad, F.mode, dact_[L.srcA], F.wa, F.ha, F.ca);
        }
        if (L.mode == 2) {
            const size_t nb = (size_t)F.W * F.H * F.cb;
            hipLaunchKernelGGL(k_stage_back_b, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, dx_, F.W, F.H,
                               F.cinPad, F.ca, dact_[L.srcB], F.cb);
        }
    }
    if (!apply) return hipGetLastError() == hipSuccess;
    ++step_;
    const float bc1 = 1.0f - powf(0.9f, (float)step_), bc2 = 1.0f - powf(0.999f, (float)step_);
    for (int li = 0; li < nl; ++li) {
        const size_t nw = layer_weights(li), nb = (size_t)pad32(Ls[li].cout);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, mw_[li], gw_[li], m1w_[li],
                           m2w_[li], nw, lr, bc1, bc2);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, mb_[li], gb_[li], m1b_[li],
                           m2b_[li], nb, lr, bc1, bc2);
    }
    if (!pack(s)) return false;
    // The forward biases are the fp32 masters.
    for (int li = 0; li < nl; ++li)
        if (hipMemcpyAsync(b_[li], mb_[li], (size_t)pad32(Ls[li].cout) * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return false;
    return hipGetLastError() == hipSuccess;
}

bool CnnNet::export_weights(std::vector<float>& blob) const {
    return training() && export_blob(mw_, mb_, blob);
}

bool CnnNet::export_grads(std::vector<float>& blob, hipStream_t s) {
    if (!training() || hipStreamSynchronize(s) != hipSuccess || !export_blob(gw_, gb_, blob)) return false;
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    for (size_t li = 0; li < Ls.size(); ++li) {
        if (hipMemset(gw_[li], 0, layer_weights((int)li) * 4) != hipSuccess) return false;
        if (hipMemset(gb_[li], 0, (size_t)pad32(Ls[li].cout) * 4) != hipSuccess) return false;
    }
    return true;
}

bool CnnNet::export_blob(const std::vector<float*>& mw, const std::vector<float*>& mb, std::vector<float>& blob) const {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    blob.clear();
    for (size_t li = 0; li < Ls.size(); ++li) {
        const CnnLayerSpec& L = Ls[li];
        const size_t nw = layer_weights((int)li);
        const int coutP = pad32(L.cout), cinP = layer_cin_pad((int)li), padA = pad32(L.cinA), cin = L.cinA + L.cinB;
        std::vector<float> w(nw), b(coutP);
        if (hipMemcpy(w.data(), mw[li], nw * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
        if (hipMemcpy(b.data(), mb[li], coutP * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
        for (int co = 0; co < L.cout; ++co)
            for (int ci = 0; ci < cin; ++ci) {
                const int cp = ci < L.cinA ? ci : padA + (ci - L.cinA);
                const int pc = packed_channel((int)li, co);
                for (int tap = 0; tap < 9; ++tap) blob.push_back(w[((size_t)tap * coutP + pc) * cinP + cp]);
            }
        for (int co = 0; co < L.cout; ++co) blob.push_back(b[packed_channel((int)li, co)]);
    }
    return true;
}

}  // namespace sr

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
    uint16_t* out;
    int relu;
};

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

template <int NT>
__global__ void __launch_bounds__(256) k_conv3x3(ConvArgs A) {
    __shared__ uint4 tile[kHalo * 4];  // 180 pixels x 32 channels of bf16
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx0 = blockIdx.x * kTW, ty0 = blockIdx.y * kTH;
    f32x4 acc[2][NT];
    for (int mt = 0; mt < 2; ++mt)
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < A.cinPad; c0 += 32) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < kHalo * 4; idx += 256) {
            const int p = idx >> 2, part = idx & 3;
            tile[idx] = fetch8(A, c0 + part * 8, ty0 - 1 + p / kHW, tx0 - 1 + p % kHW);
        }
        __syncthreads();
        const uint16_t* wc = A.w + c0 + (lane >> 4) * 8;
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 af[2];
            for (int mt = 0; mt < 2; ++mt) {
                const int p = (wave * 2 + mt + dy) * kHW + (lane & 15) + dx;
                af[mt] = __builtin_bit_cast(bf16x8, tile[p * 4 + (lane >> 4)]);
            }
            const uint16_t* wt = wc + (size_t)tap * A.coutPad * A.cinPad;
            for (int nt = 0; nt < NT; ++nt) {
                const uint4 wv = *(const uint4*)(wt + (size_t)(nt * 16 + (lane & 15)) * A.cinPad);
                const bf16x8 bfr = __builtin_bit_cast(bf16x8, wv);
                for (int mt = 0; mt < 2; ++mt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mt], bfr, acc[mt][nt], 0, 0, 0);
            }
        }
    }
    for (int mt = 0; mt < 2; ++mt) {
        const int y = ty0 + wave * 2 + mt;
        if (y >= A.H) continue;
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + (lane & 15);
            if (co >= A.cout) continue;
            const float bias = A.bias[co];
            for (int r = 0; r < 4; ++r) {
                const int x = tx0 + 4 * (lane >> 4) + r;
                if (x >= A.W) continue;
                float v = acc[mt][nt][r] + bias;
                if (A.relu) v = v > 0.0f ? v : 0.0f;
                A.out[((size_t)y * A.W + x) * A.coutPad + co] = f2bf(v);
            }
        }
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
    o[1] = make_uint4(pack2(f[8], f[9]), 0u, 0u, 0u);
    o[2] = make_uint4(0u, 0u, 0u, 0u);
    o[3] = make_uint4(0u, 0u, 0u, 0u);
}

// Softmax of the 25 logits weights the 5 x 5 neighbourhood (edge-clamped) of the noisy radiance.
__global__ void k_kpn_apply(const uint16_t* logits, const float* rad, int W, int H, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int x = i % W, y = i / W;
    float l[kTaps];
    float m = -INFINITY;
    for (int k = 0; k < kTaps; ++k) {
        l[k] = bf2f(logits[(size_t)i * 32 + k]);
        m = fmaxf(m, l[k]);
    }
    float s = 0.0f, r = 0.0f, g = 0.0f, b = 0.0f;
    for (int k = 0; k < kTaps; ++k) {
        const float e = expf(l[k] - m);
        const int yy = min(max(y + k / 5 - 2, 0), H - 1), xx = min(max(x + k % 5 - 2, 0), W - 1);
        const float* p = rad + ((size_t)yy * W + xx) * 3;
        s += e;
        r += e * p[0];
        g += e * p[1];
        b += e * p[2];
    }
    out[i * 3] = r / s;
    out[i * 3 + 1] = g / s;
    out[i * 3 + 2] = b / s;
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
        {32, 0, kTaps, 0, 17, -1, 0, false},  // 18 kernel logits
    };
    return L;
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
}

void CnnNet::release_acts() {
    if (feat_) (void)hipFree(feat_);
    feat_ = nullptr;
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
                for (int tap = 0; tap < 9; ++tap)
                    w[((size_t)tap * coutP + co) * cinP + cp] = host_f2bf(src[((size_t)co * cin + ci) * 9 + tap]);
            }
        src += (size_t)L.cout * cin * 9;
        std::vector<float> b(coutP, 0.0f);
        for (int co = 0; co < L.cout; ++co) b[co] = src[co];
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

bool CnnNet::run(const float* rad, const float* alb, const float* nrm, const float* dep, uint32_t W, uint32_t H,
                 float* out, hipStream_t s) {
    if (!loaded_ || !ensure(W, H)) return false;
    const int n = (int)(W * H);
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, s, rad, alb, nrm, dep, (int)W, (int)H, feat_);
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    for (size_t li = 0; li < Ls.size(); ++li) {
        const CnnLayerSpec& L = Ls[li];
        ConvArgs A;
        memset(&A, 0, sizeof(A));
        uint32_t w, h;
        level_dims(W, H, L.level, w, h);
        const int srcLevel = L.mode == 1 ? L.level - 1 : (L.mode == 2 ? L.level + 1 : L.level);
        uint32_t wa, ha;
        level_dims(W, H, srcLevel, wa, ha);
        A.a = L.srcA < 0 ? feat_ : act_[L.srcA];
        A.ca = L.srcA < 0 ? 32 : pad32(Ls[L.srcA].cout);
        A.b = L.srcB < 0 ? nullptr : act_[L.srcB];
        A.cb = L.srcB < 0 ? 0 : pad32(Ls[L.srcB].cout);
        A.wa = (int)wa;
        A.ha = (int)ha;
        A.mode = L.mode;
        A.W = (int)w;
        A.H = (int)h;
        A.cinPad = cin_pad(L, A.ca, A.cb);
        A.coutPad = pad32(L.cout);
        A.cout = L.cout;
        A.w = w_[li];
        A.bias = b_[li];
        A.out = act_[li];
        A.relu = L.relu ? 1 : 0;
        const dim3 grid((w + kTW - 1) / kTW, (h + kTH - 1) / kTH);
        switch ((L.cout + 15) / 16) {
            case 2: hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, s, A); break;
            case 3: hipLaunchKernelGGL(k_conv3x3<3>, grid, dim3(256), 0, s, A); break;
            case 4: hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(256), 0, s, A); break;
            case 6: hipLaunchKernelGGL(k_conv3x3<6>, grid, dim3(256), 0, s, A); break;
            case 8: hipLaunchKernelGGL(k_conv3x3<8>, grid, dim3(256), 0, s, A); break;
            default: return false;
        }
    }
    hipLaunchKernelGGL(k_kpn_apply, dim3((n + 255) / 256), dim3(256), 0, s, act_.back(), rad, (int)W, (int)H, out);
    return hipGetLastError() == hipSuccess;
}

bool CnnNet::read_layer(int layer, std::vector<float>& out, uint32_t& c, uint32_t& w, uint32_t& h) const {
    const std::vector<CnnLayerSpec>& Ls = cnn_layers();
    if (!feat_ || layer < -1 || layer >= (int)Ls.size()) return false;
    const int level = layer < 0 ? 0 : Ls[layer].level;
    level_dims(W_, H_, level, w, h);
    c = layer < 0 ? kFeat : Ls[layer].cout;
    const int stride = layer < 0 ? 32 : pad32(Ls[layer].cout);
    std::vector<uint16_t> raw((size_t)w * h * stride);
    if (hipMemcpy(raw.data(), layer < 0 ? feat_ : act_[layer], raw.size() * 2, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    out.resize((size_t)w * h * c);
    for (size_t p = 0; p < (size_t)w * h; ++p)
        for (uint32_t k = 0; k < c; ++k) out[p * c + k] = host_bf2f(raw[p * stride + k]);
    return true;
}

}  // namespace sr

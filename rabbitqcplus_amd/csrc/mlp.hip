// Training of the learned seam's classifier on the GPU: an MLP 16 -> 64 -> 64 -> 1
// (ReLU, ReLU, logit) with binary cross-entropy on the logit, hand-written forward and
// backward kernels behind a C++ forward/backward Function, and Adam.
//
// The classifier replaces the forest of the reference's correctAnchorClf seam
// (include/corrector.hpp:1448-1472, include/forest.hpp); the training itself has no
// counterpart in the reference (its forests are trained offline by scikit-learn).
//
// Parameter layout (float, host and device): W1[64][16], b1[64], W2[64][64], b2[64],
// W3[64], b3 -- CARE_MLP_PARAMS values.
//
// Kernels work on tiles of 64 rows with 256 threads: thread t owns row t % 64 and the
// hidden units [16 * (t / 64), +16), so at each step of a sum the 64 lanes of a wave read
// 64 different activation rows (a padded stride, no bank conflicts) and one broadcast
// weight. Weight gradients are summed per tile, and the tiles' partial sums are reduced in
// tile order, so a gradient does not depend on scheduling.

#include <hip/hip_runtime.h>

#include "care_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int kIn = 16;
constexpr int kH = 64;
constexpr int kP = kH * kIn + kH + kH * kH + kH + kH + 1;
constexpr int kOffW1 = 0, kOffB1 = kH * kIn, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;
constexpr int kRows = 64;          // rows per tile
constexpr int kThreads = 256;
constexpr int kPad = kH + 1;       // padded row stride of the activation tiles
static_assert(kP == CARE_MLP_PARAMS, "parameter count");

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b){
        if(b <= bytes) return hipSuccess;
        if(p){ hipError_t e = hipFree(p); if(e != hipSuccess) return e; p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, b);
        if(e == hipSuccess) bytes = b;
        return e;
    }
    template<class T> T* as() const { return static_cast<T*>(p); }
    ~DevBuf(){ if(p) (void)hipFree(p); }
};

// ---------------------------------------------------------------- kernels

// Forward of one tile: z[r], and the post-ReLU activations H1, H2 kept for backward
// (not written when H1out is null: inference only).
__global__ void __launch_bounds__(kThreads)
k_forward(const float* __restrict__ P, const float* __restrict__ X, int n, float* __restrict__ z,
          float* __restrict__ H1out, float* __restrict__ H2out){
    __shared__ float sX[kRows][kIn + 1];
    __shared__ float sW1[kH][kIn + 1];
    __shared__ float sW2[kH][kPad];
    __shared__ float sH1[kRows][kPad];
    __shared__ float sH2[kRows][kPad];
    __shared__ float sB[2 * kH + kH + 1];    // b1, b2, W3, b3
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * kRows;
    for(int q = t; q < kRows * kIn; q += kThreads){
        const int r = q / kIn, c = q - r * kIn;
        sX[r][c] = r0 + r < n ? X[size_t(r0 + r) * kIn + c] : 0.0f;
    }
    for(int q = t; q < kH * kIn; q += kThreads) sW1[q / kIn][q % kIn] = P[kOffW1 + q];
    for(int q = t; q < kH * kH; q += kThreads) sW2[q / kH][q % kH] = P[kOffW2 + q];
    for(int q = t; q < kH; q += kThreads){
        sB[q] = P[kOffB1 + q]; sB[kH + q] = P[kOffB2 + q]; sB[2 * kH + q] = P[kOffW3 + q];
    }
    if(t == 0) sB[3 * kH] = P[kOffB3];
    __syncthreads();
    const int r = t % kRows;
    const int j0 = (t / kRows) * 16;
    float acc[16];
    #pragma unroll
    for(int u = 0; u < 16; u++) acc[u] = sB[j0 + u];
    for(int k = 0; k < kIn; k++){
        const float x = sX[r][k];
        #pragma unroll
        for(int u = 0; u < 16; u++) acc[u] += sW1[j0 + u][k] * x;
    }
    #pragma unroll
    for(int u = 0; u < 16; u++) sH1[r][j0 + u] = acc[u] > 0.0f ? acc[u] : 0.0f;
    __syncthreads();
    #pragma unroll
    for(int u = 0; u < 16; u++) acc[u] = sB[kH + j0 + u];
    for(int k = 0; k < kH; k++){
        const float h = sH1[r][k];
        #pragma unroll
        for(int u = 0; u < 16; u++) acc[u] += sW2[j0 + u][k] * h;
    }
    #pragma unroll
    for(int u = 0; u < 16; u++) sH2[r][j0 + u] = acc[u] > 0.0f ? acc[u] : 0.0f;
    __syncthreads();
    if(H1out && r0 + r < n){
        float* h1 = H1out + size_t(r0 + r) * kH + j0;
        float* h2 = H2out + size_t(r0 + r) * kH + j0;
        #pragma unroll
        for(int u = 0; u < 16; u++){ h1[u] = sH1[r][j0 + u]; h2[u] = sH2[r][j0 + u]; }
    }
    if(t < kRows && r0 + t < n){
        float s = sB[3 * kH];
        for(int j = 0; j < kH; j++) s += sB[2 * kH + j] * sH2[t][j];
        z[r0 + t] = s;
    }
}

// Binary cross-entropy on logits: dz = (sigmoid(z) - y) / n, and per-tile loss sums.
__global__ void __launch_bounds__(kThreads)
k_bce(const float* __restrict__ z, const float* __restrict__ y, int n, float* __restrict__ dz,
      float* __restrict__ lossPart){
    __shared__ float red[kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    float l = 0.0f;
    if(i < n){
        const float v = z[i];
        const float yy = y[i];
        l = fmaxf(v, 0.0f) - yy * v + log1pf(expf(-fabsf(v)));
        dz[i] = (1.0f / (1.0f + expf(-v)) - yy) / float(n);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for(int s = kThreads / 2; s > 0; s >>= 1){
        if(threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if(threadIdx.x == 0) lossPart[blockIdx.x] = red[0];
}

// Backward of one tile: the tile's partial gradient (all kP parameters) into part[tile].
__global__ void __launch_bounds__(kThreads)
k_backward(const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ H1in,
           const float* __restrict__ H2in, const float* __restrict__ dz, int n, float* __restrict__ part){
    __shared__ float sX[kRows][kIn + 1];
    __shared__ float sA1[kRows][kPad];    // H1, then dH1
    __shared__ float sD2[kRows][kPad];    // H2, then dH2
    __shared__ float sW2[kH][kPad];
    __shared__ float sW3[kH];
    __shared__ float sDz[kRows];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * kRows;
    for(int q = t; q < kRows * kIn; q += kThreads){
        const int r = q / kIn, c = q - r * kIn;
        sX[r][c] = r0 + r < n ? X[size_t(r0 + r) * kIn + c] : 0.0f;
    }
    for(int q = t; q < kRows * kH; q += kThreads){
        const int r = q / kH, c = q - r * kH;
        const bool ok = r0 + r < n;
        sA1[r][c] = ok ? H1in[size_t(r0 + r) * kH + c] : 0.0f;
        sD2[r][c] = ok ? H2in[size_t(r0 + r) * kH + c] : 0.0f;
    }
    for(int q = t; q < kH * kH; q += kThreads) sW2[q / kH][q % kH] = P[kOffW2 + q];
    if(t < kH) sW3[t] = P[kOffW3 + t];
    if(t < kRows) sDz[t] = r0 + t < n ? dz[r0 + t] : 0.0f;
    __syncthreads();
    float* g = part + size_t(blockIdx.x) * kP;
    // output layer: gW3[j] = sum_r dz[r] H2[r][j], gb3 = sum_r dz[r]
    if(t < kH){
        float s = 0.0f;
        for(int r = 0; r < kRows; r++) s += sDz[r] * sD2[r][t];
        g[kOffW3 + t] = s;
    }else if(t == kH){
        float s = 0.0f;
        for(int r = 0; r < kRows; r++) s += sDz[r];
        g[kOffB3] = s;
    }
    __syncthreads();
    // dH2 = dz * W3 where H2 > 0
    for(int q = t; q < kRows * kH; q += kThreads){
        const int r = q / kH, c = q - r * kH;
        sD2[r][c] = sD2[r][c] > 0.0f ? sDz[r] * sW3[c] : 0.0f;
    }
    __syncthreads();
    // dH1[r][k] = sum_j dH2[r][j] W2[j][k] where H1 > 0 (held in registers until H1 is
    // no longer needed)
    float dh1[16];
    const int rr = t % kRows;
    const int kk0 = (t / kRows) * 16;
    {
        #pragma unroll
        for(int u = 0; u < 16; u++) dh1[u] = 0.0f;
        for(int j = 0; j < kH; j++){
            const float d = sD2[rr][j];
            #pragma unroll
            for(int u = 0; u < 16; u++) dh1[u] += d * sW2[j][kk0 + u];
        }
        #pragma unroll
        for(int u = 0; u < 16; u++) dh1[u] = sA1[rr][kk0 + u] > 0.0f ? dh1[u] : 0.0f;
    }
    // gW2[j][k] = sum_r dH2[r][j] H1[r][k]; thread owns j = t % 64, k in [16 * (t / 64), +16)
    {
        const int j = t % kH;
        const int k0 = (t / kH) * 16;
        float acc[16];
        #pragma unroll
        for(int u = 0; u < 16; u++) acc[u] = 0.0f;
        for(int r = 0; r < kRows; r++){
            const float d = sD2[r][j];
            #pragma unroll
            for(int u = 0; u < 16; u++) acc[u] += d * sA1[r][k0 + u];
        }
        #pragma unroll
        for(int u = 0; u < 16; u++) g[kOffW2 + j * kH + k0 + u] = acc[u];
    }
    // gb2[j] = sum_r dH2[r][j]
    if(t < kH){
        float s = 0.0f;
        for(int r = 0; r < kRows; r++) s += sD2[r][t];
        g[kOffB2 + t] = s;
    }
    __syncthreads();
    #pragma unroll
    for(int u = 0; u < 16; u++) sA1[rr][kk0 + u] = dh1[u];
    __syncthreads();
    // gb1[j] = sum_r dH1[r][j]
    if(t < kH){
        float s = 0.0f;
        for(int r = 0; r < kRows; r++) s += sA1[r][t];
        g[kOffB1 + t] = s;
    }
    // gW1[j][m] = sum_r dH1[r][j] X[r][m]; thread owns j = t % 64, m in [4 * (t / 64), +4)
    {
        const int j = t % kH;
        const int m0 = (t / kH) * 4;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for(int r = 0; r < kRows; r++){
            const float d = sA1[r][j];
            #pragma unroll
            for(int u = 0; u < 4; u++) acc[u] += d * sX[r][m0 + u];
        }
        #pragma unroll
        for(int u = 0; u < 4; u++) g[kOffW1 + j * kIn + m0 + u] = acc[u];
    }
}

// grad[p] = sum over tiles, in tile order.
__global__ void k_reduce_tiles(const float* __restrict__ part, int tiles, float* __restrict__ grad){
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= kP) return;
    float s = 0.0f;
    for(int b = 0; b < tiles; b++) s += part[size_t(b) * kP + p];
    grad[p] = s;
}

// Sum of the per-block loss parts divided by n, added to *acc with weight w.
__global__ void k_loss_sum(const float* __restrict__ lossPart, int parts, int n, float w, float* __restrict__ out,
                           float* __restrict__ acc){
    if(threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for(int b = 0; b < parts; b++) s += lossPart[b];
    const float l = s / float(n);
    if(out) *out = l;
    if(acc) *acc += w * l;
}

// Adam (Kingma & Ba 2015) with bias correction: c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float beta1, float beta2, float eps, float c1, float c2){
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= kP) return;
    const float gg = g[p];
    const float mm = beta1 * m[p] + (1.0f - beta1) * gg;
    const float vv = beta2 * v[p] + (1.0f - beta2) * gg * gg;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ y, const uint32_t* __restrict__ perm,
                              int n, float* __restrict__ Xb, float* __restrict__ yb){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n * kIn) return;
    const int r = i / kIn, c = i - r * kIn;
    const uint32_t src = perm[r];
    Xb[i] = X[size_t(src) * kIn + c];
    if(c == 0) yb[r] = y[src];
}

__global__ void k_fill(float* v, int n, float x){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) v[i] = x;
}

}  // namespace

// Device scratch of care_mlp_forward_static, the allocation-free inference call: the
// input rows and the logits of one call (at most kStaticRows rows) live in this
// file-scope array, which every GPU the library runs on holds until the process ends.
// Host code never passes the array's name to a kernel or a copy (in host code it is a
// stand-in, not a device address): the device address of each GPU's copy is taken once
// per device, at care_mlp_create with that device current, and kept in gScratch.
constexpr size_t kStaticRows = 65536;
constexpr size_t kStaticXBytes = kStaticRows * kIn * sizeof(float);
constexpr size_t kStaticBytes = kStaticXBytes + kStaticRows * sizeof(float);
static __device__ __align__(256) uint8_t scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), resolved once.
hipError_t scratchAddress(int device, void** out){
    if(device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if(!gScratch[device]){
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if(e != hipSuccess) return e;
        gScratch[device] = ptr;
    }
    *out = gScratch[device];
    return hipSuccess;
}

// The C++ forward/backward Function: forward(X) -> z saves what backward needs;
// backward(dz) -> dL/dparams given dL/dz. Device pointers, one stream.
class MlpFunction {
public:
    hipError_t forward(const float* P, const float* X, int n, float* z, hipStream_t st){
        hipError_t e;
        if((e = h1_.ensure(size_t(std::max(n, 1)) * kH * 4)) != hipSuccess) return e;
        if((e = h2_.ensure(size_t(std::max(n, 1)) * kH * 4)) != hipSuccess) return e;
        X_ = X; n_ = n;
        if(n == 0) return hipSuccess;
        k_forward<<<(n + kRows - 1) / kRows, kThreads, 0, st>>>(P, X, n, z, h1_.as<float>(), h2_.as<float>());
        return hipGetLastError();
    }
    hipError_t backward(const float* P, const float* dz, float* grad, hipStream_t st){
        hipError_t e;
        const int tiles = std::max(1, (n_ + kRows - 1) / kRows);
        if((e = part_.ensure(size_t(tiles) * kP * 4)) != hipSuccess) return e;
        if(n_ == 0){
            k_fill<<<(kP + 255) / 256, 256, 0, st>>>(grad, kP, 0.0f);
            return hipGetLastError();
        }
        k_backward<<<tiles, kThreads, 0, st>>>(P, X_, h1_.as<float>(), h2_.as<float>(), dz, n_, part_.as<float>());
        if((e = hipGetLastError()) != hipSuccess) return e;
        k_reduce_tiles<<<(kP + 255) / 256, 256, 0, st>>>(part_.as<float>(), tiles, grad);
        return hipGetLastError();
    }
private:
    DevBuf h1_, h2_, part_;
    const float* X_ = nullptr;
    int n_ = 0;
};

struct care_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    MlpFunction fn;
    DevBuf params, grad, m, v, z, dz, lossPart, lossOut, X, y, Xall, yall, Xb, yb, perm, epochLoss;
    size_t nData = 0;
    uint64_t step = 0;
    void* scratch = nullptr;       // this device's address of the static scratch array

    int fail(int code, const char* fmt, ...){
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define MCHECK(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess){ \
    return h->fail(CARE_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); } } while(0)

namespace {

// forward + BCE + backward on rows already on the device; loss (device float) to lossOut
int lossGradDevice(care_mlp* h, const float* X, const float* y, int n){
    hipStream_t st = h->stream;
    MCHECK(h->z.ensure(size_t(std::max(n, 1)) * 4));
    MCHECK(h->dz.ensure(size_t(std::max(n, 1)) * 4));
    const int parts = std::max(1, (n + kThreads - 1) / kThreads);
    MCHECK(h->lossPart.ensure(size_t(parts) * 4));
    MCHECK(h->fn.forward(h->params.as<float>(), X, n, h->z.as<float>(), st));
    if(n > 0){
        k_bce<<<parts, kThreads, 0, st>>>(h->z.as<float>(), y, n, h->dz.as<float>(), h->lossPart.as<float>());
        MCHECK(hipGetLastError());
    }else{
        k_fill<<<1, 64, 0, st>>>(h->lossPart.as<float>(), 1, 0.0f);
    }
    MCHECK(h->fn.backward(h->params.as<float>(), h->dz.as<float>(), h->grad.as<float>(), st));
    return CARE_OK;
}

}  // namespace

// Entry points run their bodies through guarded(): no C++ exception crosses the
// C-ABI (a failed host allocation is CARE_ENOMEM, any other exception CARE_EINVAL).
template<class F>
int guarded(care_mlp* h, F&& f){
    const char* what = "unexpected C++ exception";
    int code = CARE_EINVAL;
    try{
        return f();
    }catch(const std::bad_alloc&){
        what = "host allocation failed";
        code = CARE_ENOMEM;
    }catch(const std::exception& e){
        what = e.what();
    }catch(...){
    }
    if(h){
        try{ h->err = what; }catch(...){}
    }
    return code;
}

extern "C" {

static int care_mlp_create_body(int device, care_mlp** out){
    if(!out) return CARE_EINVAL;
    *out = nullptr;
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return CARE_ENODEV;
    auto* h = new care_mlp;
    h->device = device;
    if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess){
        delete h;
        return CARE_EHIP;
    }
    if(h->params.ensure(kP * 4) != hipSuccess || h->grad.ensure(kP * 4) != hipSuccess || h->m.ensure(kP * 4) != hipSuccess ||
       h->v.ensure(kP * 4) != hipSuccess || h->lossOut.ensure(16) != hipSuccess){
        delete h;
        return CARE_ENOMEM;
    }
    k_fill<<<(kP + 255) / 256, 256, 0, h->stream>>>(h->params.as<float>(), kP, 0.0f);
    k_fill<<<(kP + 255) / 256, 256, 0, h->stream>>>(h->m.as<float>(), kP, 0.0f);
    k_fill<<<(kP + 255) / 256, 256, 0, h->stream>>>(h->v.as<float>(), kP, 0.0f);
    if(hipStreamSynchronize(h->stream) != hipSuccess){ delete h; return CARE_EHIP; }
    // the device address of the static scratch of care_mlp_forward_static (device current)
    if(scratchAddress(device, &h->scratch) != hipSuccess){ care_mlp_destroy(h); return CARE_EHIP; }
    *out = h;
    return CARE_OK;
}

void care_mlp_destroy(care_mlp* h){
    if(!h) return;
    (void)hipSetDevice(h->device);
    if(h->stream) (void)hipStreamSynchronize(h->stream);
    if(h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* care_mlp_last_error(const care_mlp* h){ return h ? h->err.c_str() : "null handle"; }

static int care_mlp_set_params_body(care_mlp* h, const float* params, size_t count){
    if(!h || !params) return CARE_EINVAL;
    if(count != size_t(kP)) return h->fail(CARE_EINVAL, "expected %d parameters, got %zu", kP, count);
    MCHECK(hipSetDevice(h->device));
    MCHECK(hipMemcpy(h->params.p, params, kP * 4, hipMemcpyHostToDevice));
    return CARE_OK;
}

static int care_mlp_get_params_body(care_mlp* h, float* params, size_t count){
    if(!h || !params) return CARE_EINVAL;
    if(count != size_t(kP)) return h->fail(CARE_EINVAL, "expected %d parameters, got %zu", kP, count);
    MCHECK(hipSetDevice(h->device));
    MCHECK(hipStreamSynchronize(h->stream));
    MCHECK(hipMemcpy(params, h->params.p, kP * 4, hipMemcpyDeviceToHost));
    return CARE_OK;
}

static int care_mlp_reset_optimizer_body(care_mlp* h){
    if(!h) return CARE_EINVAL;
    MCHECK(hipSetDevice(h->device));
    k_fill<<<(kP + 255) / 256, 256, 0, h->stream>>>(h->m.as<float>(), kP, 0.0f);
    k_fill<<<(kP + 255) / 256, 256, 0, h->stream>>>(h->v.as<float>(), kP, 0.0f);
    MCHECK(hipGetLastError());
    h->step = 0;
    return CARE_OK;
}

static int care_mlp_forward_body(care_mlp* h, const float* X, size_t n, float* z){
    if(!h || (n && (!X || !z))) return CARE_EINVAL;
    if(n > size_t(INT32_MAX)) return h->fail(CARE_EINVAL, "too many rows");
    MCHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    MCHECK(h->X.ensure(std::max<size_t>(n, 1) * kIn * 4));
    MCHECK(h->z.ensure(std::max<size_t>(n, 1) * 4));
    if(n) MCHECK(hipMemcpyAsync(h->X.p, X, n * kIn * 4, hipMemcpyHostToDevice, st));
    MCHECK(h->fn.forward(h->params.as<float>(), h->X.as<float>(), int(n), h->z.as<float>(), st));
    if(n) MCHECK(hipMemcpyAsync(z, h->z.p, n * 4, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    return CARE_OK;
}

static int care_mlp_forward_static_body(care_mlp* h, const float* X, size_t n, float* z){
    if(!h || (n && (!X || !z))) return CARE_EINVAL;
    if(n > kStaticRows) return h->fail(CARE_EINVAL, "care_mlp_forward_static takes at most %zu rows per call",
                                       kStaticRows);
    if(!h->scratch) return h->fail(CARE_ESTATE, "static scratch not resolved for device %d", h->device);
    if(n == 0) return CARE_OK;
    MCHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    float* sx = static_cast<float*>(h->scratch);
    float* sz = reinterpret_cast<float*>(static_cast<uint8_t*>(h->scratch) + kStaticXBytes);
    MCHECK(hipMemcpyAsync(sx, X, n * kIn * 4, hipMemcpyHostToDevice, st));
    k_forward<<<int((n + kRows - 1) / kRows), kThreads, 0, st>>>(h->params.as<float>(), sx, int(n), sz, nullptr, nullptr);
    MCHECK(hipGetLastError());
    MCHECK(hipMemcpyAsync(z, sz, n * 4, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    return CARE_OK;
}

static int care_mlp_loss_grad_body(care_mlp* h, const float* X, const float* y, size_t n, float* loss, float* grad){
    if(!h || (n && (!X || !y))) return CARE_EINVAL;
    if(n > size_t(INT32_MAX)) return h->fail(CARE_EINVAL, "too many rows");
    MCHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    MCHECK(h->X.ensure(std::max<size_t>(n, 1) * kIn * 4));
    MCHECK(h->y.ensure(std::max<size_t>(n, 1) * 4));
    if(n){
        MCHECK(hipMemcpyAsync(h->X.p, X, n * kIn * 4, hipMemcpyHostToDevice, st));
        MCHECK(hipMemcpyAsync(h->y.p, y, n * 4, hipMemcpyHostToDevice, st));
    }
    const int rc = lossGradDevice(h, h->X.as<float>(), h->y.as<float>(), int(n));
    if(rc != CARE_OK) return rc;
    const int parts = std::max(1, int((n + kThreads - 1) / kThreads));
    k_loss_sum<<<1, 64, 0, st>>>(h->lossPart.as<float>(), parts, int(std::max<size_t>(n, 1)), 0.0f, h->lossOut.as<float>(), nullptr);
    MCHECK(hipGetLastError());
    if(loss) MCHECK(hipMemcpyAsync(loss, h->lossOut.p, 4, hipMemcpyDeviceToHost, st));
    if(grad) MCHECK(hipMemcpyAsync(grad, h->grad.p, kP * 4, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    return CARE_OK;
}

static int care_mlp_set_data_body(care_mlp* h, const float* X, const float* y, size_t n){
    if(!h || (n && (!X || !y))) return CARE_EINVAL;
    if(n > size_t(INT32_MAX)) return h->fail(CARE_EINVAL, "too many rows");
    MCHECK(hipSetDevice(h->device));
    MCHECK(h->Xall.ensure(std::max<size_t>(n, 1) * kIn * 4));
    MCHECK(h->yall.ensure(std::max<size_t>(n, 1) * 4));
    if(n){
        MCHECK(hipMemcpy(h->Xall.p, X, n * kIn * 4, hipMemcpyHostToDevice));
        MCHECK(hipMemcpy(h->yall.p, y, n * 4, hipMemcpyHostToDevice));
    }
    h->nData = n;
    return CARE_OK;
}

static int care_mlp_train_body(care_mlp* h, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss){
    if(!h || epochs < 0 || batch <= 0) return CARE_EINVAL;
    if(h->nData == 0) return h->fail(CARE_ESTATE, "no training data (care_mlp_set_data)");
    MCHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const size_t n = h->nData;
    const int bsz = int(std::min<size_t>(size_t(batch), n));
    const size_t batches = (n + size_t(bsz) - 1) / size_t(bsz);
    MCHECK(h->perm.ensure(n * 4));
    MCHECK(h->Xb.ensure(size_t(bsz) * kIn * 4));
    MCHECK(h->yb.ensure(size_t(bsz) * 4));
    MCHECK(h->epochLoss.ensure(std::max(epochs, 1) * 4));
    k_fill<<<(std::max(epochs, 1) + 255) / 256, 256, 0, st>>>(h->epochLoss.as<float>(), std::max(epochs, 1), 0.0f);
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> perm(n);
    for(int ep = 0; ep < epochs; ep++){
        std::iota(perm.begin(), perm.end(), 0u);
        std::shuffle(perm.begin(), perm.end(), rng);
        MCHECK(hipMemcpyAsync(h->perm.p, perm.data(), n * 4, hipMemcpyHostToDevice, st));
        for(size_t b = 0; b < batches; b++){
            const int rows = int(std::min<size_t>(size_t(bsz), n - b * bsz));
            k_gather_rows<<<(rows * kIn + 255) / 256, 256, 0, st>>>(h->Xall.as<float>(), h->yall.as<float>(),
                                                                   h->perm.as<uint32_t>() + b * bsz, rows,
                                                                   h->Xb.as<float>(), h->yb.as<float>());
            MCHECK(hipGetLastError());
            const int rc = lossGradDevice(h, h->Xb.as<float>(), h->yb.as<float>(), rows);
            if(rc != CARE_OK) return rc;
            const int parts = (rows + kThreads - 1) / kThreads;
            k_loss_sum<<<1, 64, 0, st>>>(h->lossPart.as<float>(), parts, rows, 1.0f / float(batches), nullptr,
                                         h->epochLoss.as<float>() + ep);
            h->step++;
            const double t = double(h->step);
            const float c1 = float(1.0 - std::pow(double(beta1), t));
            const float c2 = float(1.0 - std::pow(double(beta2), t));
            k_adam<<<(kP + 255) / 256, 256, 0, st>>>(h->params.as<float>(), h->grad.as<float>(), h->m.as<float>(),
                                                     h->v.as<float>(), lr, beta1, beta2, eps, c1, c2);
            MCHECK(hipGetLastError());
        }
        // the permutation buffer is rewritten next epoch: wait for this epoch's gathers
        MCHECK(hipStreamSynchronize(st));
    }
    if(epoch_loss && epochs > 0) MCHECK(hipMemcpy(epoch_loss, h->epochLoss.p, size_t(epochs) * 4, hipMemcpyDeviceToHost));
    MCHECK(hipStreamSynchronize(st));
    return CARE_OK;
}


// Exported entry points: each runs its body under guarded().
int care_mlp_create(int device, care_mlp** out){
    return guarded(nullptr, [&]() { return care_mlp_create_body(device, out); });
}

int care_mlp_set_params(care_mlp* h, const float* params, size_t count){
    return guarded(h, [&]() { return care_mlp_set_params_body(h, params, count); });
}

int care_mlp_get_params(care_mlp* h, float* params, size_t count){
    return guarded(h, [&]() { return care_mlp_get_params_body(h, params, count); });
}

int care_mlp_reset_optimizer(care_mlp* h){
    return guarded(h, [&]() { return care_mlp_reset_optimizer_body(h); });
}

int care_mlp_forward(care_mlp* h, const float* X, size_t n, float* z){
    return guarded(h, [&]() { return care_mlp_forward_body(h, X, n, z); });
}

int care_mlp_loss_grad(care_mlp* h, const float* X, const float* y, size_t n, float* loss, float* grad){
    return guarded(h, [&]() { return care_mlp_loss_grad_body(h, X, y, n, loss, grad); });
}

int care_mlp_forward_static(care_mlp* h, const float* X, size_t n, float* z){
    return guarded(h, [&]() { return care_mlp_forward_static_body(h, X, n, z); });
}

int care_mlp_set_data(care_mlp* h, const float* X, const float* y, size_t n){
    return guarded(h, [&]() { return care_mlp_set_data_body(h, X, y, n); });
}

int care_mlp_train(care_mlp* h, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss){
    return guarded(h, [&]() { return care_mlp_train_body(h, epochs, batch, lr, seed, epoch_loss); });
}

}  // extern "C"

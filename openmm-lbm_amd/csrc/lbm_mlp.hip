/*
 * The closure network (include/lbm_amd_closure.h) on gfx950: 87 -> 64 -> 64 -> 3 with SiLU, trained and evaluated by
 * FP16 MFMA kernels with FP32 accumulation.
 *
 * Wave layout of v_mfma_f32_16x16x16_f16 (64 lanes; lane l, row = l % 16, quad = 4 (l / 16)):
 *   A operand (M x K):  lane holds A[row][quad .. quad + 3]
 *   B operand (K x N):  lane holds B[quad .. quad + 3][row]
 *   C / D     (M x N):  lane holds D[quad .. quad + 3][row]
 * The network runs transposed, Z^T = W^T X^T, with the 16 particles of a tile as N: the D tile of a layer's outputs
 * 16t .. 16t + 15 is then, element for element, the B operand of the next layer's K step t, so activations stay in
 * registers from the input to the output; the weights are A operands held in registers by every wave for the whole
 * kernel.  Backward, dH^T = W dZ^T has the same property.  Per tile of 16 particles: 24 + 16 + 4 MFMAs forward,
 * 4 + 16 more backward.
 *
 * Training (lbm_mlp_train): per minibatch, mlpTrainForward (forward, error, backward through the activations; writes
 * the FP16 activations and errors row by row), mlpWeightGrad (dW = P^T Q over chunks of 256 particles, one 16 x 16 tile
 * of one matrix per wave, partial sums per chunk), mlpUpdate (the chunks summed in a fixed order, Adam on the FP32
 * master parameters, new FP16 operands).  Every sum has a fixed order: training is reproducible run to run.
 */
#include "lbm_mlp.h"
#include <algorithm>
#include <cmath>
#include <mutex>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>

namespace lbm {
void setLastError(const std::string& message);
}

using namespace lbm;

namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;                 // waves per work-group of the network kernels
constexpr int CHUNK = 256;               // particles per partial sum of the weight gradients
constexpr int GRAD_JOBS = 24 + 16 + 4;   // 16 x 16 tiles of dW1, dW2, dW3

__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float siluf(float z) {
    return z/(1.0f + __expf(-z));
}

__device__ __forceinline__ float dsiluf(float z) {
    float s = 1.0f/(1.0f + __expf(-z));
    return s*(1.0f + z*(1.0f - s));
}

__device__ __forceinline__ half4_t load4(const _Float16* p) {
    return *(const half4_t*) p;
}

__device__ __forceinline__ void store4(_Float16* p, half4_t v) {
    *(half4_t*) p = v;
}

struct ForwardWeights {
    half4_t w1[4][6], w2[4][4], w3[4];
    float4_t b1[4], b2[4], b3;
};

__device__ __forceinline__ void loadForward(const MlpDevice* __restrict__ m, int row, int quad, ForwardWeights& w) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int s = 0; s < 6; s++)
            w.w1[t][s] = load4(m->w1t + (16*t + row)*MLP_IN_PAD + 16*s + quad);
#pragma unroll
        for (int s = 0; s < 4; s++)
            w.w2[t][s] = load4(m->w2t + (16*t + row)*MLP_H + 16*s + quad);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            w.b1[t][r] = m->b1[16*t + quad + r];
            w.b2[t][r] = m->b2[16*t + quad + r];
        }
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
        w.w3[s] = load4(m->w3t + row*MLP_H + 16*s + quad);
#pragma unroll
    for (int r = 0; r < 4; r++)
        w.b3[r] = m->b3[quad + r];
}

struct Activations {
    float4_t z1[4], z2[4], y;
    half4_t h1[4], h2[4];
};

/* The forward pass of one tile: x[s] is the B operand of input columns 16s .. 16s + 15. */
__device__ __forceinline__ void forwardTile(const ForwardWeights& w, const half4_t x[6], Activations& a) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float4_t acc = w.b1[t];
#pragma unroll
        for (int s = 0; s < 6; s++)
            acc = mfma16(w.w1[t][s], x[s], acc);
        a.z1[t] = acc;
#pragma unroll
        for (int r = 0; r < 4; r++)
            a.h1[t][r] = (_Float16) siluf(acc[r]);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float4_t acc = w.b2[t];
#pragma unroll
        for (int s = 0; s < 4; s++)
            acc = mfma16(w.w2[t][s], a.h1[s], acc);
        a.z2[t] = acc;
#pragma unroll
        for (int r = 0; r < 4; r++)
            a.h2[t][r] = (_Float16) siluf(acc[r]);
    }
    float4_t acc = w.b3;
#pragma unroll
    for (int s = 0; s < 4; s++)
        acc = mfma16(w.w3[s], a.h2[s], acc);
    a.y = acc;
}

/* Inference: persistent waves, each loading the weights once and looping over tiles of 16 particles. */
__global__ void __launch_bounds__(64*WAVES) mlpInfer(const MlpDevice* __restrict__ m, const _Float16* __restrict__ X,
        int n, float outInv, float* __restrict__ du) {
    int lane = threadIdx.x & 63;
    int wave = blockIdx.x*WAVES + (threadIdx.x >> 6);
    int waves = gridDim.x*WAVES;
    int tiles = (n + 15)/16;
    if (wave >= tiles)
        return;
    int row = lane & 15, quad = (lane >> 4)*4;
    ForwardWeights w;
    loadForward(m, row, quad, w);
    for (int tile = wave; tile < tiles; tile += waves) {
        int b = tile*16 + row;
        bool valid = b < n;
        half4_t x[6];
#pragma unroll
        for (int s = 0; s < 6; s++)
            x[s] = valid ? load4(X + (size_t) b*MLP_IN_PAD + 16*s + quad) : half4_t{0, 0, 0, 0};
        Activations a;
        forwardTile(w, x, a);
        if (lane < 16 && valid)
            for (int r = 0; r < MLP_OUT; r++)
                du[3*(size_t) b + r] = a.y[r]*outInv;
    }
}

/*
 * Training, forward and backward through the activations, one wave per tile of 16 rows of the minibatch (rows[b] is the
 * dataset row of minibatch row b; rows at or beyond count are padding and get zero errors).  Per row b it stores
 * H1, H2 (post-activation), dZ1, dZ2 (errors at the pre-activations) and dY = (2/3)(y - t) as FP16; the loss sum
 * goes to lossSum (a diagnostic).
 */
__global__ void __launch_bounds__(64*WAVES) mlpTrainForward(const MlpDevice* __restrict__ m,
        const _Float16* __restrict__ X, const float* __restrict__ T, const int* __restrict__ rows, int count, int tiles,
        _Float16* __restrict__ H1, _Float16* __restrict__ H2, _Float16* __restrict__ dZ1, _Float16* __restrict__ dZ2,
        _Float16* __restrict__ dY, double* __restrict__ lossSum) {
    int lane = threadIdx.x & 63;
    int tile = blockIdx.x*WAVES + (threadIdx.x >> 6);
    if (tile >= tiles)
        return;
    int row = lane & 15, quad = (lane >> 4)*4;
    ForwardWeights w;
    loadForward(m, row, quad, w);
    int b = tile*16 + row;
    bool valid = b < count;
    int src = rows[b];
    half4_t x[6];
#pragma unroll
    for (int s = 0; s < 6; s++)
        x[s] = load4(X + (size_t) src*MLP_IN_PAD + 16*s + quad);
    Activations a;
    forwardTile(w, x, a);
    // Output error (lanes 0-15 hold outputs 0-3 of their row; output 3 and up are padding).
    float4_t dy = {0, 0, 0, 0};
    float e = 0;
    if (quad == 0 && valid)
        for (int r = 0; r < MLP_OUT; r++) {
            float d = a.y[r] - T[3*(size_t) src + r];
            dy[r] = (2.0f/3.0f)*d;
            e += d*d;
        }
    half4_t dyb;
#pragma unroll
    for (int r = 0; r < 4; r++)
        dyb[r] = (_Float16) dy[r];
    // Backward: dH2^T = W3 dY^T, dZ2 = dH2 silu'(Z2), dH1^T = W2 dZ2^T, dZ1 = dH1 silu'(Z1).
    half4_t dz2[4], dz1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        half4_t w3n = load4(m->w3 + (16*t + row)*MLP_OUT_PAD + quad);
        float4_t acc = mfma16(w3n, dyb, float4_t{0, 0, 0, 0});
#pragma unroll
        for (int r = 0; r < 4; r++)
            dz2[t][r] = (_Float16) (acc[r]*dsiluf(a.z2[t][r]));
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float4_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++)
            acc = mfma16(load4(m->w2 + (16*t + row)*MLP_H + 16*s + quad), dz2[s], acc);
#pragma unroll
        for (int r = 0; r < 4; r++)
            dz1[t][r] = (_Float16) (acc[r]*dsiluf(a.z1[t][r]));
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        store4(H1 + (size_t) b*MLP_H + 16*t + quad, a.h1[t]);
        store4(H2 + (size_t) b*MLP_H + 16*t + quad, a.h2[t]);
        store4(dZ1 + (size_t) b*MLP_H + 16*t + quad, dz1[t]);
        store4(dZ2 + (size_t) b*MLP_H + 16*t + quad, dz2[t]);
    }
    store4(dY + (size_t) b*MLP_OUT_PAD + quad, dyb);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        e += __shfl_xor(e, off);
    if (lane == 0)
        atomicAdd(lossSum, (double) e/3.0);
}

/*
 * Weight gradients: wave (chunk, job) computes one 16 x 16 tile of dW = P^T Q over the 256 minibatch rows of its chunk,
 * P the layer input (X through rows, H1, H2) and Q the layer error (dZ1, dZ2, dY), and, for the first row tile, the
 * bias gradient (column sums of Q).  Written to partials[chunk][parameter].
 */
__global__ void __launch_bounds__(64*WAVES) mlpWeightGrad(const _Float16* __restrict__ X, const int* __restrict__ rows,
        const _Float16* __restrict__ H1, const _Float16* __restrict__ H2, const _Float16* __restrict__ dZ1,
        const _Float16* __restrict__ dZ2, const _Float16* __restrict__ dY, int chunks, float* __restrict__ partials) {
    int lane = threadIdx.x & 63;
    int wave = blockIdx.x*WAVES + (threadIdx.x >> 6);
    int chunk = wave/GRAD_JOBS, job = wave % GRAD_JOBS;
    if (chunk >= chunks)
        return;
    int row = lane & 15, quad = (lane >> 4)*4;
    const _Float16* P;
    const _Float16* Q;
    int pw, qw, ti, to, offW, offB, width;
    bool gather = false;
    if (job < 24) {
        P = X; pw = MLP_IN_PAD; Q = dZ1; qw = MLP_H; ti = job/4; to = job % 4;
        offW = MLP_OFF_W1; offB = MLP_OFF_B1; width = MLP_H; gather = true;
    }
    else if (job < 40) {
        P = H1; pw = MLP_H; Q = dZ2; qw = MLP_H; ti = (job - 24)/4; to = (job - 24) % 4;
        offW = MLP_OFF_W2; offB = MLP_OFF_B2; width = MLP_H;
    }
    else {
        P = H2; pw = MLP_H; Q = dY; qw = MLP_OUT_PAD; ti = job - 40; to = 0;
        offW = MLP_OFF_W3; offB = MLP_OFF_B3; width = MLP_OUT_PAD;
    }
    float4_t acc = {0, 0, 0, 0};
    float bsum = 0;
    for (int k = 0; k < CHUNK/16; k++) {
        int b0 = chunk*CHUNK + 16*k + quad;
        half4_t pa, qb;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int b = b0 + e;
            int prow = gather ? rows[b] : b;
            pa[e] = P[(size_t) prow*pw + 16*ti + row];
            qb[e] = Q[(size_t) b*qw + 16*to + row];
            bsum += (float) qb[e];
        }
        acc = mfma16(pa, qb, acc);
    }
    float* out = partials + (size_t) chunk*MLP_PARAMS;
#pragma unroll
    for (int r = 0; r < 4; r++)
        out[offW + (16*ti + quad + r)*width + 16*to + row] = acc[r];
    if (ti == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (lane < 16)
            out[offB + 16*to + lane] = bsum;
    }
}

/* The FP16 operands and FP32 biases of parameter p. */
__device__ __forceinline__ void storeOperand(MlpDevice* d, int p, float w) {
    _Float16 h = (_Float16) w;
    if (p < MLP_OFF_B1) {
        int i = p/MLP_H, o = p % MLP_H;
        d->w1t[o*MLP_IN_PAD + i] = h;
    }
    else if (p < MLP_OFF_W2)
        d->b1[p - MLP_OFF_B1] = w;
    else if (p < MLP_OFF_B2) {
        int q = p - MLP_OFF_W2, i = q/MLP_H, o = q % MLP_H;
        d->w2t[o*MLP_H + i] = h;
        d->w2[i*MLP_H + o] = h;
    }
    else if (p < MLP_OFF_W3)
        d->b2[p - MLP_OFF_B2] = w;
    else if (p < MLP_OFF_B3) {
        int q = p - MLP_OFF_W3, i = q/MLP_OUT_PAD, o = q % MLP_OUT_PAD;
        d->w3t[o*MLP_H + i] = h;
        d->w3[i*MLP_OUT_PAD + o] = h;
    }
    else
        d->b3[p - MLP_OFF_B3] = w;
}

__global__ void mlpPack(const float* __restrict__ master, MlpDevice* __restrict__ d) {
    int p = blockIdx.x*blockDim.x + threadIdx.x;
    if (p < MLP_PARAMS)
        storeOperand(d, p, master[p]);
}

/* Sum of the chunk partials (fixed order) / count -> gradient; with apply, one Adam step (beta 0.9, 0.999, eps 1e-8;
   c1, c2 the bias corrections 1/(1 - beta^t)) and the new FP16 operands. */
__global__ void mlpUpdate(float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
        const float* __restrict__ partials, int chunks, float invCount, float lr, float c1, float c2,
        MlpDevice* __restrict__ d, float* __restrict__ gradOut, int apply) {
    int p = blockIdx.x*blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS)
        return;
    float g = 0;
    for (int c = 0; c < chunks; c++)
        g += partials[(size_t) c*MLP_PARAMS + p];
    g *= invCount;
    if (gradOut)
        gradOut[p] = g;
    if (!apply)
        return;
    float mm = 0.9f*m1[p] + 0.1f*g;
    float vv = 0.999f*m2[p] + 0.001f*g*g;
    m1[p] = mm;
    m2[p] = vv;
    float w = master[p] - lr*(mm*c1)/(sqrtf(vv*c2) + 1e-8f);
    master[p] = w;
    storeOperand(d, p, w);
}

/* Raw pairs -> normalized FP16 inputs [n][96] and normalized FP32 targets [n][3]. */
__global__ void mlpNormalize(const float* __restrict__ raw, const float* __restrict__ rawTargets, int64_t n,
        ClosureNorm norm, _Float16* __restrict__ X, float* __restrict__ T) {
    int64_t idx = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
    if (idx < n*MLP_IN_PAD) {
        int64_t b = idx/MLP_IN_PAD;
        int c = (int) (idx - b*MLP_IN_PAD);
        float v = 0;
        if (c < 3)
            v = raw[b*MLP_IN + c];
        else if (c < 84)
            v = raw[b*MLP_IN + c]*norm.uScale;
        else if (c < MLP_IN)
            v = raw[b*MLP_IN + c]*norm.vScale;
        X[idx] = (_Float16) v;
    }
    if (rawTargets != nullptr && idx < n*MLP_OUT)
        T[idx] = rawTargets[idx]*norm.outScale;
}

struct MlpError : public std::runtime_error {
    int status;
    MlpError(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw MlpError(LBM_ERROR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* deviceAlloc(size_t count) {
    void* p = nullptr;
    hipCheck(hipMalloc(&p, std::max<size_t>(count, 1)*sizeof(T)), "hipMalloc");
    return (T*) p;
}

unsigned blocksFor(int64_t n, int threads) {
    return (unsigned) ((n + threads - 1)/threads);
}

}  // namespace

/*
 * The device working memory of lbm_mlp_forward_noalloc: raw inputs, normalized FP16 inputs and outputs of up to
 * LBM_MLP_NOALLOC_MAX_ROWS rows (9.0 MB), one copy per GPU for the life of the process.  Every call on a GPU shares
 * that GPU's copy; nothing in it has to start at zero (each call writes every element it reads).
 */
namespace {
constexpr size_t NOALLOC_RAW_BYTES = (size_t) LBM_MLP_NOALLOC_MAX_ROWS*MLP_IN*sizeof(float);
constexpr size_t NOALLOC_X_BYTES = (size_t) LBM_MLP_NOALLOC_MAX_ROWS*MLP_IN_PAD*sizeof(_Float16);
constexpr size_t NOALLOC_OUT_BYTES = (size_t) LBM_MLP_NOALLOC_MAX_ROWS*MLP_OUT*sizeof(float);
static_assert(NOALLOC_RAW_BYTES % 256 == 0 && NOALLOC_X_BYTES % 256 == 0, "256-byte aligned parts");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_RAW_BYTES + NOALLOC_X_BYTES + NOALLOC_OUT_BYTES];

namespace {
std::mutex scratchMutex;
std::vector<void*> scratchByDevice;

/* The device address of `scratch` on the current device (the host-side name is only a stand-in), looked up once per
   device with that device current. */
void* scratchOnCurrentDevice() {
    int device = 0;
    hipCheck(hipGetDevice(&device), "hipGetDevice");
    std::lock_guard<std::mutex> lock(scratchMutex);
    if ((int) scratchByDevice.size() <= device)
        scratchByDevice.resize(device + 1, nullptr);
    if (scratchByDevice[device] == nullptr) {
        void* ptr = nullptr;
        hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        scratchByDevice[device] = ptr;
    }
    return scratchByDevice[device];
}
}  // namespace

struct lbm_mlp {
    hipStream_t stream = nullptr;
    void* scratch = nullptr;          // this GPU's no-allocation working memory (scratchOnCurrentDevice)
    MlpDevice* dev = nullptr;
    float* master = nullptr;
    float* mom1 = nullptr;
    float* mom2 = nullptr;
    lbm_mlp_normalization norm{1.0, 1.0, 1.0, 1e30, 1e30};
    int64_t adamStep = 0;
};

namespace lbm {

const MlpDevice* mlpDeviceWeights(const lbm_mlp* mlp) {
    return mlp->dev;
}

ClosureNorm mlpClosureNorm(const lbm_mlp* mlp) {
    ClosureNorm n;
    n.uScale = (float) mlp->norm.u_scale;
    n.vScale = (float) mlp->norm.v_scale;
    n.outScale = (float) mlp->norm.out_scale;
    n.uBound = (float) std::min(mlp->norm.u_bound, 1e30);
    n.vBound = (float) std::min(mlp->norm.v_bound, 1e30);
    return n;
}

void launchMlpInfer(const MlpDevice* weights, const _Float16* X, int n, float outScale, float* du, hipStream_t stream) {
    if (n <= 0)
        return;
    int tiles = (n + 15)/16;
    // Persistent waves, one per SIMD of the 256 compute units: each loads the weights once.
    int waves = std::min(tiles, 1024);
    mlpInfer<<<blocksFor(waves, WAVES), 64*WAVES, 0, stream>>>(weights, X, n, 1.0f/outScale, du);
}

}  // namespace lbm

namespace {

/* Minibatch buffers of the training kernels, for up to `rows` rows (a multiple of CHUNK). */
struct TrainBuffers {
    int capacity = 0;
    _Float16 *H1 = nullptr, *H2 = nullptr, *dZ1 = nullptr, *dZ2 = nullptr, *dY = nullptr;
    float* partials = nullptr;
    double* loss = nullptr;
    void allocate(int rows) {
        capacity = rows;
        H1 = deviceAlloc<_Float16>((size_t) rows*MLP_H);
        H2 = deviceAlloc<_Float16>((size_t) rows*MLP_H);
        dZ1 = deviceAlloc<_Float16>((size_t) rows*MLP_H);
        dZ2 = deviceAlloc<_Float16>((size_t) rows*MLP_H);
        dY = deviceAlloc<_Float16>((size_t) rows*MLP_OUT_PAD);
        partials = deviceAlloc<float>((size_t) (rows/CHUNK)*MLP_PARAMS);
        loss = deviceAlloc<double>(1);
    }
    ~TrainBuffers() {
        void* all[] = {H1, H2, dZ1, dZ2, dY, partials, loss};
        for (void* p : all)
            if (p)
                (void) hipFree(p);
    }
};

/* Dataset on the device: normalized inputs and targets. */
struct Dataset {
    int64_t n = 0;
    _Float16* X = nullptr;
    float* T = nullptr;
    Dataset(lbm_mlp* m, const float* features, const float* targets, int64_t count) : n(count) {
        float* raw = deviceAlloc<float>((size_t) n*MLP_IN);
        float* rawT = targets ? deviceAlloc<float>((size_t) n*MLP_OUT) : nullptr;
        X = deviceAlloc<_Float16>((size_t) n*MLP_IN_PAD);
        T = deviceAlloc<float>((size_t) n*MLP_OUT);
        hipError_t e = hipMemcpyAsync(raw, features, (size_t) n*MLP_IN*sizeof(float), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess && rawT)
            e = hipMemcpyAsync(rawT, targets, (size_t) n*MLP_OUT*sizeof(float), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) {
            mlpNormalize<<<blocksFor(n*MLP_IN_PAD, 256), 256, 0, m->stream>>>(raw, rawT, n, mlpClosureNorm(m), X, T);
            e = hipStreamSynchronize(m->stream);
        }
        (void) hipStreamSynchronize(m->stream);
        (void) hipFree(raw);
        if (rawT)
            (void) hipFree(rawT);
        if (e != hipSuccess) {
            (void) hipFree(X);
            (void) hipFree(T);
            X = nullptr;
            T = nullptr;
            hipCheck(e, "dataset upload");
        }
    }
    ~Dataset() {
        if (X)
            (void) hipFree(X);
        if (T)
            (void) hipFree(T);
    }
};

/* One minibatch: rows[0 .. count) of the dataset (rows padded with valid indices to a multiple of CHUNK). */
void minibatch(lbm_mlp* m, const Dataset& data, const int* rows, int count, TrainBuffers& buf, bool apply, float lr,
        float* gradOut) {
    int chunks = (count + CHUNK - 1)/CHUNK;
    int tiles = chunks*(CHUNK/16);
    mlpTrainForward<<<blocksFor(tiles, WAVES), 64*WAVES, 0, m->stream>>>(m->dev, data.X, data.T, rows, count, tiles,
            buf.H1, buf.H2, buf.dZ1, buf.dZ2, buf.dY, buf.loss);
    mlpWeightGrad<<<blocksFor((int64_t) chunks*GRAD_JOBS, WAVES), 64*WAVES, 0, m->stream>>>(data.X, rows, buf.H1, buf.H2,
            buf.dZ1, buf.dZ2, buf.dY, chunks, buf.partials);
    float c1 = 1.0f, c2 = 1.0f;
    if (apply) {
        m->adamStep++;
        c1 = (float) (1.0/(1.0 - std::pow(0.9, (double) m->adamStep)));
        c2 = (float) (1.0/(1.0 - std::pow(0.999, (double) m->adamStep)));
    }
    mlpUpdate<<<blocksFor(MLP_PARAMS, 256), 256, 0, m->stream>>>(m->master, m->mom1, m->mom2, buf.partials, chunks,
            1.0f/count, lr, c1, c2, m->dev, gradOut, apply ? 1 : 0);
    hipCheck(hipGetLastError(), "training launch");
}

/* Scales and bounds from raw pairs: inverse RMS of the fluid velocities, of the bead velocity and of the targets;
   bounds = largest normalized magnitude of each group. */
lbm_mlp_normalization normalizationOf(const float* features, const float* targets, int64_t n) {
    double su = 0, sv = 0, st = 0, mu = 0, mv = 0;
    for (int64_t b = 0; b < n; b++) {
        const float* f = features + b*MLP_IN;
        for (int c = 3; c < 84; c++) {
            su += (double) f[c]*f[c];
            mu = std::max(mu, (double) std::fabs(f[c]));
        }
        for (int c = 84; c < MLP_IN; c++) {
            sv += (double) f[c]*f[c];
            mv = std::max(mv, (double) std::fabs(f[c]));
        }
        for (int c = 0; c < MLP_OUT; c++)
            st += (double) targets[b*MLP_OUT + c]*targets[b*MLP_OUT + c];
    }
    lbm_mlp_normalization norm;
    double ru = std::sqrt(su/(81.0*n)), rv = std::sqrt(sv/(3.0*n)), rt = std::sqrt(st/(3.0*n));
    norm.u_scale = ru > 0 ? 1.0/ru : 1.0;
    norm.v_scale = rv > 0 ? 1.0/rv : 1.0;
    norm.out_scale = rt > 0 ? 1.0/rt : 1.0;
    norm.u_bound = mu*norm.u_scale;
    norm.v_bound = mv*norm.v_scale;
    return norm;
}

int failMlp(const std::exception& e, int status) {
    setLastError(e.what());
    return status;
}

void destroyMlp(lbm_mlp* m) {
    if (m->stream)
        (void) hipStreamSynchronize(m->stream);
    void* all[] = {m->dev, m->master, m->mom1, m->mom2};
    for (void* p : all)
        if (p)
            (void) hipFree(p);
    if (m->stream)
        (void) hipStreamDestroy(m->stream);
    delete m;
}

void uploadParameters(lbm_mlp* m, const float* params) {
    hipCheck(hipMemcpyAsync(m->master, params, MLP_PARAMS*sizeof(float), hipMemcpyHostToDevice, m->stream), "upload");
    hipCheck(hipMemsetAsync(m->mom1, 0, MLP_PARAMS*sizeof(float), m->stream), "memset");
    hipCheck(hipMemsetAsync(m->mom2, 0, MLP_PARAMS*sizeof(float), m->stream), "memset");
    mlpPack<<<blocksFor(MLP_PARAMS, 256), 256, 0, m->stream>>>(m->master, m->dev);
    hipCheck(hipGetLastError(), "pack");
    hipCheck(hipStreamSynchronize(m->stream), "sync");
    m->adamStep = 0;
}

/* The padding entries of a parameter vector must be zero (they multiply padded inputs or feed padded outputs). */
void checkPadding(const float* params) {
    for (int i = MLP_IN; i < MLP_IN_PAD; i++)
        for (int o = 0; o < MLP_H; o++)
            if (params[MLP_OFF_W1 + i*MLP_H + o] != 0.0f)
                throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_set_parameters: rows 87-95 of W1 (input padding) must be zero");
    for (int i = 0; i < MLP_H; i++)
        for (int o = MLP_OUT; o < MLP_OUT_PAD; o++)
            if (params[MLP_OFF_W3 + i*MLP_OUT_PAD + o] != 0.0f)
                throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_set_parameters: columns 3-15 of W3 (output padding) must be zero");
    for (int o = MLP_OUT; o < MLP_OUT_PAD; o++)
        if (params[MLP_OFF_B3 + o] != 0.0f)
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_set_parameters: entries 3-15 of b3 (output padding) must be zero");
}

}  // namespace

#define MLP_TRY(...) \
    try { __VA_ARGS__; setLastError(""); return LBM_OK; } \
    catch (const MlpError& e) { return failMlp(e, e.status); } \
    catch (const std::exception& e) { return failMlp(e, LBM_ERROR_RUNTIME); }

extern "C" {

int lbm_mlp_create(uint64_t seed, lbm_mlp** out) {
    if (!out) {
        setLastError("lbm_mlp_create: null argument");
        return LBM_ERROR_INVALID;
    }
    lbm_mlp* m = new lbm_mlp();
    try {
        hipCheck(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
        m->scratch = scratchOnCurrentDevice();
        m->dev = deviceAlloc<MlpDevice>(1);
        m->master = deviceAlloc<float>(MLP_PARAMS);
        m->mom1 = deviceAlloc<float>(MLP_PARAMS);
        m->mom2 = deviceAlloc<float>(MLP_PARAMS);
        std::vector<float> p(MLP_PARAMS, 0.0f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> normal(0.0, 1.0);
        for (int i = 0; i < MLP_IN; i++)
            for (int o = 0; o < MLP_H; o++)
                p[MLP_OFF_W1 + i*MLP_H + o] = (float) (normal(rng)/std::sqrt((double) MLP_IN));
        for (int i = 0; i < MLP_H; i++)
            for (int o = 0; o < MLP_H; o++)
                p[MLP_OFF_W2 + i*MLP_H + o] = (float) (normal(rng)/std::sqrt((double) MLP_H));
        for (int i = 0; i < MLP_H; i++)
            for (int o = 0; o < MLP_OUT; o++)
                p[MLP_OFF_W3 + i*MLP_OUT_PAD + o] = (float) (normal(rng)/std::sqrt((double) MLP_H));
        hipCheck(hipMemsetAsync(m->dev, 0, sizeof(MlpDevice), m->stream), "memset");
        uploadParameters(m, p.data());
    }
    catch (const MlpError& e) {
        destroyMlp(m);
        return failMlp(e, e.status);
    }
    catch (const std::exception& e) {
        destroyMlp(m);
        return failMlp(e, LBM_ERROR_RUNTIME);
    }
    *out = m;
    setLastError("");
    return LBM_OK;
}

int lbm_mlp_destroy(lbm_mlp* m) {
    if (m)
        destroyMlp(m);
    return LBM_OK;
}

int lbm_mlp_get_parameters(lbm_mlp* m, float* params, size_t count) {
    MLP_TRY({
        if (count != (size_t) MLP_PARAMS)
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_get_parameters: count must be LBM_MLP_NUM_PARAMS");
        hipCheck(hipStreamSynchronize(m->stream), "sync");
        hipCheck(hipMemcpy(params, m->master, MLP_PARAMS*sizeof(float), hipMemcpyDeviceToHost), "download");
    })
}

int lbm_mlp_set_parameters(lbm_mlp* m, const float* params, size_t count) {
    MLP_TRY({
        if (count != (size_t) MLP_PARAMS)
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_set_parameters: count must be LBM_MLP_NUM_PARAMS");
        checkPadding(params);
        uploadParameters(m, params);
    })
}

int lbm_mlp_get_normalization(lbm_mlp* m, lbm_mlp_normalization* norm) {
    MLP_TRY(*norm = m->norm)
}

int lbm_mlp_set_normalization(lbm_mlp* m, const lbm_mlp_normalization* norm) {
    MLP_TRY({
        if (!(norm->u_scale > 0 && norm->v_scale > 0 && norm->out_scale > 0) || !(norm->u_bound >= 0 && norm->v_bound >= 0))
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_set_normalization: scales must be positive and bounds non-negative");
        m->norm = *norm;
    })
}

int lbm_mlp_train(lbm_mlp* m, const float* features, const float* targets, int64_t n, int epochs, int batch,
        double lr, double decay, uint64_t seed, double* epochLoss) {
    MLP_TRY({
        if (n <= 0 || n > (int64_t) 0x7FFFFFFF - (1 << 20) || epochs <= 0 || batch < 16 || batch > (1 << 20) || !(lr > 0))
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_train: invalid n, epochs, batch or learning rate");
        m->norm = normalizationOf(features, targets, n);
        Dataset data(m, features, targets, n);
        int padded = (batch + CHUNK - 1)/CHUNK*CHUNK;
        TrainBuffers buf;
        buf.allocate(padded);
        int* order = deviceAlloc<int>((size_t) n + padded);
        std::vector<int> perm((size_t) n + padded, 0);
        std::mt19937_64 rng(seed);
        try {
            for (int epoch = 0; epoch < epochs; epoch++) {
                std::iota(perm.begin(), perm.begin() + n, 0);
                std::shuffle(perm.begin(), perm.begin() + n, rng);
                hipCheck(hipMemcpyAsync(order, perm.data(), perm.size()*sizeof(int), hipMemcpyHostToDevice, m->stream), "upload");
                hipCheck(hipMemsetAsync(buf.loss, 0, sizeof(double), m->stream), "memset");
                float rate = (float) (lr*std::pow(decay, (double) epoch));
                for (int64_t b0 = 0; b0 < n; b0 += batch) {
                    int count = (int) std::min<int64_t>(batch, n - b0);
                    minibatch(m, data, order + b0, count, buf, true, rate, nullptr);
                }
                double loss = 0;
                hipCheck(hipMemcpyAsync(&loss, buf.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream), "download");
                hipCheck(hipStreamSynchronize(m->stream), "sync");
                if (epochLoss)
                    epochLoss[epoch] = loss/(double) n;
            }
        }
        catch (...) {
            (void) hipStreamSynchronize(m->stream);
            (void) hipFree(order);
            throw;
        }
        (void) hipFree(order);
    })
}

int lbm_mlp_forward(lbm_mlp* m, const float* features, int64_t n, float* out) {
    MLP_TRY({
        if (n <= 0 || n > 0x7FFFFFFF)
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_forward: invalid n");
        Dataset data(m, features, nullptr, n);
        float* du = deviceAlloc<float>((size_t) n*MLP_OUT);
        launchMlpInfer(m->dev, data.X, (int) n, (float) m->norm.out_scale, du, m->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, du, (size_t) n*MLP_OUT*sizeof(float), hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(m->stream);
        (void) hipStreamSynchronize(m->stream);
        (void) hipFree(du);
        hipCheck(e, "forward");
    })
}

int lbm_mlp_forward_noalloc(lbm_mlp* m, const float* features, int64_t n, float* out) {
    MLP_TRY({
        if (n <= 0 || n > LBM_MLP_NOALLOC_MAX_ROWS) {
            std::stringstream msg;
            msg << "lbm_mlp_forward_noalloc: n = " << n << " is outside [1, LBM_MLP_NOALLOC_MAX_ROWS = "
                << LBM_MLP_NOALLOC_MAX_ROWS << "]";
            throw MlpError(LBM_ERROR_INVALID, msg.str());
        }
        char* base = (char*) m->scratch;
        float* raw = (float*) base;
        _Float16* X = (_Float16*) (base + NOALLOC_RAW_BYTES);
        float* du = (float*) (base + NOALLOC_RAW_BYTES + NOALLOC_X_BYTES);
        hipCheck(hipMemcpyAsync(raw, features, (size_t) n*MLP_IN*sizeof(float), hipMemcpyHostToDevice, m->stream),
                "forward upload");
        mlpNormalize<<<blocksFor(n*MLP_IN_PAD, 256), 256, 0, m->stream>>>(raw, nullptr, n, mlpClosureNorm(m), X,
                nullptr);
        launchMlpInfer(m->dev, X, (int) n, (float) m->norm.out_scale, du, m->stream);
        hipCheck(hipGetLastError(), "forward launch");
        hipCheck(hipMemcpyAsync(out, du, (size_t) n*MLP_OUT*sizeof(float), hipMemcpyDeviceToHost, m->stream),
                "forward download");
        hipCheck(hipStreamSynchronize(m->stream), "forward");
    })
}

int lbm_mlp_gradients(lbm_mlp* m, const float* features, const float* targets, int64_t n, float* grad, double* loss) {
    MLP_TRY({
        if (n <= 0 || n > (1 << 20))
            throw MlpError(LBM_ERROR_INVALID, "lbm_mlp_gradients: n must be in [1, 2^20]");
        Dataset data(m, features, targets, n);
        int padded = (int) ((n + CHUNK - 1)/CHUNK*CHUNK);
        TrainBuffers buf;
        buf.allocate(padded);
        std::vector<int> rows(padded, 0);
        std::iota(rows.begin(), rows.begin() + n, 0);
        int* dRows = deviceAlloc<int>(padded);
        float* dGrad = deviceAlloc<float>(MLP_PARAMS);
        double lossSum = 0;
        hipError_t e = hipMemcpyAsync(dRows, rows.data(), padded*sizeof(int), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(buf.loss, 0, sizeof(double), m->stream);
        if (e == hipSuccess) {
            minibatch(m, data, dRows, (int) n, buf, false, 0.0f, dGrad);
            e = hipMemcpyAsync(grad, dGrad, MLP_PARAMS*sizeof(float), hipMemcpyDeviceToHost, m->stream);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(&lossSum, buf.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(m->stream);
        (void) hipStreamSynchronize(m->stream);
        (void) hipFree(dRows);
        (void) hipFree(dGrad);
        hipCheck(e, "gradients");
        if (loss)
            *loss = lossSum/(double) n;
    })
}

}  // extern "C"

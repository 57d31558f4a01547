// Learned edge proposal network: per-edge features (24) -> 64 -> 64 -> 1
// logit, ReLU hidden layers.  No counterpart in the reference (its edge
// sampler is the edge tree alone, edge_tree.cpp:511-725).
//
// Inference runs on the matrix cores: bf16 weights, fp32 accumulation with
// v_mfma_f32_16x16x16_bf16.  A wave evaluates 16 edges.  Each layer computes
// H^T = W^T X^T, so that the accumulator layout of one layer (lane l holds
// rows 4*(l/16)+j, column l%16) is exactly the B-operand layout of the next
// layer's k step, and activations never leave registers.
//
// Training keeps fp32 master weights: a workgroup takes a minibatch of 64
// edges, runs forward and backward in fp32 from LDS and adds its parameter
// gradients to a global buffer; an Adam step then updates the master weights
// and refreshes the bf16 transposed copies used by inference.
// Loss: weighted squared error of the logit against a target log-weight.
#include "../../include/delta_ray.h"
#include "dr_math.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#define MLP_IN 24
#define MLP_INP 32  // input width padded to two 16-wide k steps
#define MLP_H 64
// fp32 parameter layout: W1 [in][h], b1 [h], W2 [h][h] (in x out), b2 [h], w3 [h], b3
#define P_W1 0
#define P_B1 (P_W1 + MLP_IN * MLP_H)
#define P_W2 (P_B1 + MLP_H)
#define P_B2 (P_W2 + MLP_H * MLP_H)
#define P_W3 (P_B2 + MLP_H)
#define P_B3 (P_W3 + MLP_H)
#define P_COUNT (P_B3 + 1)
#define MB 64  // training minibatch per workgroup

static_assert(P_COUNT == DR_MLP_PARAMS, "parameter count");

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

extern "C" dr_status dr_internal_fail(dr_status s, const char *msg);

DR_INLINE unsigned short to_bf16(float f) {
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    unsigned int r = ((u >> 16) & 1u) + 0x7fffu;
    return (unsigned short)((u + r) >> 16);
}

// --------------------------------------------------------- inference ----
// W1T: [64][32] bf16 (hidden x padded input), W2T: [64][64] bf16 (out x in)
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ X, long long n,
                                                   const unsigned short *__restrict__ W1T,
                                                   const unsigned short *__restrict__ W2T,
                                                   const float *__restrict__ P, float *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const long long base = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16;
    if (base >= n) return;
    const int col = lane & 15, grp = lane >> 4;
    const long long e = base + col;
    // layer 1: B = X^T, two k steps of 16 input features
    v4s b1[2];
    for (int ks = 0; ks < 2; ks++) {
        for (int j = 0; j < 4; j++) {
            int f = ks * 16 + 4 * grp + j;
            float x = (e < n && f < MLP_IN) ? X[e * MLP_IN + f] : 0.f;
            b1[ks][j] = (short)to_bf16(x);
        }
    }
    v4s h1[4];  // B operand of layer 2, one per k step (= m tile of layer 1)
    for (int mt = 0; mt < 4; mt++) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 2; ks++) {
            v4s a;
            for (int j = 0; j < 4; j++) a[j] = (short)W1T[(mt * 16 + col) * MLP_INP + ks * 16 + 4 * grp + j];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1[ks], acc, 0, 0, 0);
        }
        for (int j = 0; j < 4; j++) {
            float h = fmaxf(acc[j] + P[P_B1 + mt * 16 + 4 * grp + j], 0.f);
            h1[mt][j] = (short)to_bf16(h);
        }
    }
    float part = 0.f;
    for (int mt = 0; mt < 4; mt++) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ks++) {
            v4s a;
            for (int j = 0; j < 4; j++) a[j] = (short)W2T[(mt * 16 + col) * MLP_H + ks * 16 + 4 * grp + j];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h1[ks], acc, 0, 0, 0);
        }
        for (int j = 0; j < 4; j++) {
            int k = mt * 16 + 4 * grp + j;
            float h = fmaxf(acc[j] + P[P_B2 + k], 0.f);
            part += P[P_W3 + k] * h;
        }
    }
    // the four lane groups hold the partial sums of one edge
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    if (grp == 0 && e < n) out[e] = part + P[P_B3];
}

// ---------------------------------------------------------- training ----
// idx: minibatch edge indices (MB per workgroup); grad: fp32, accumulated
__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ X, const float *__restrict__ Y,
                                                  const float *__restrict__ Wt, const int *__restrict__ idx,
                                                  int nbatch, const float *__restrict__ P, float *grad,
                                                  float *loss, float inv_total) {
    __shared__ float x[MB][MLP_IN];
    __shared__ float z1[MB][MLP_H];
    __shared__ float z2[MB][MLP_H];
    __shared__ float dz[MB][MLP_H];
    __shared__ float dout[MB];
    const int tid = threadIdx.x;
    const int b0 = blockIdx.x * MB;
    for (int i = tid; i < MB * MLP_IN; i += blockDim.x) {
        int r = i / MLP_IN, f = i % MLP_IN;
        int e = b0 + r < nbatch ? idx[b0 + r] : -1;
        x[r][f] = e >= 0 ? X[(long long)e * MLP_IN + f] : 0.f;
    }
    __syncthreads();
    for (int i = tid; i < MB * MLP_H; i += blockDim.x) {
        int r = i / MLP_H, k = i % MLP_H;
        float s = P[P_B1 + k];
        for (int f = 0; f < MLP_IN; f++) s += x[r][f] * P[P_W1 + f * MLP_H + k];
        z1[r][k] = s;
    }
    __syncthreads();
    for (int i = tid; i < MB * MLP_H; i += blockDim.x) {
        int r = i / MLP_H, k = i % MLP_H;
        float s = P[P_B2 + k];
        for (int j = 0; j < MLP_H; j++) s += fmaxf(z1[r][j], 0.f) * P[P_W2 + j * MLP_H + k];
        z2[r][k] = s;
    }
    __syncthreads();
    if (tid < MB) {
        int e = b0 + tid < nbatch ? idx[b0 + tid] : -1;
        float d = 0.f;
        if (e >= 0) {
            float o = P[P_B3];
            for (int k = 0; k < MLP_H; k++) o += fmaxf(z2[tid][k], 0.f) * P[P_W3 + k];
            float w = Wt[e], r = o - Y[e];
            d = 2.f * w * r * inv_total;
            atomicAdd(loss, w * r * r * inv_total);
        }
        dout[tid] = d;
    }
    __syncthreads();
    // output layer
    for (int k = tid; k < MLP_H; k += blockDim.x) {
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += dout[r] * fmaxf(z2[r][k], 0.f);
        atomicAdd(grad + P_W3 + k, g);
    }
    if (tid == 0) {
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += dout[r];
        atomicAdd(grad + P_B3, g);
    }
    for (int i = tid; i < MB * MLP_H; i += blockDim.x) {
        int r = i / MLP_H, k = i % MLP_H;
        dz[r][k] = z2[r][k] > 0.f ? dout[r] * P[P_W3 + k] : 0.f;
    }
    __syncthreads();
    // second layer
    for (int i = tid; i < MLP_H * MLP_H; i += blockDim.x) {
        int j = i / MLP_H, k = i % MLP_H;
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += fmaxf(z1[r][j], 0.f) * dz[r][k];
        if (g != 0.f) atomicAdd(grad + P_W2 + i, g);
    }
    for (int k = tid; k < MLP_H; k += blockDim.x) {
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += dz[r][k];
        atomicAdd(grad + P_B2 + k, g);
    }
    __syncthreads();
    // dz1 = (W2 dz2) * relu'(z1), into z2's storage
    for (int i = tid; i < MB * MLP_H; i += blockDim.x) {
        int r = i / MLP_H, j = i % MLP_H;
        float s = 0.f;
        for (int k = 0; k < MLP_H; k++) s += P[P_W2 + j * MLP_H + k] * dz[r][k];
        z2[r][j] = z1[r][j] > 0.f ? s : 0.f;
    }
    __syncthreads();
    // first layer
    for (int i = tid; i < MLP_IN * MLP_H; i += blockDim.x) {
        int f = i / MLP_H, k = i % MLP_H;
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += x[r][f] * z2[r][k];
        if (g != 0.f) atomicAdd(grad + P_W1 + i, g);
    }
    for (int k = tid; k < MLP_H; k += blockDim.x) {
        float g = 0.f;
        for (int r = 0; r < MB; r++) g += z2[r][k];
        atomicAdd(grad + P_B1 + k, g);
    }
}

// Adam (Kingma & Ba 2015) on the fp32 master weights; then the bf16
// transposed copies for inference.  grad is cleared for the next step.
__global__ void k_mlp_adam(float *P, float *grad, float *m, float *v, int t, float lr, unsigned short *W1T,
                           unsigned short *W2T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P_COUNT) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        float g = grad[i];
        float mi = b1 * m[i] + (1.f - b1) * g;
        float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
        P[i] -= lr * mh / (sqrtf(vh) + eps);
        grad[i] = 0.f;
    }
}

__global__ void k_mlp_pack(const float *P, unsigned short *W1T, unsigned short *W2T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MLP_H * MLP_INP) {
        int k = i / MLP_INP, f = i % MLP_INP;
        W1T[i] = to_bf16(f < MLP_IN ? P[P_W1 + f * MLP_H + k] : 0.f);
    }
    if (i < MLP_H * MLP_H) {
        int k = i / MLP_H, j = i % MLP_H;  // out k, in j
        W2T[i] = to_bf16(P[P_W2 + j * MLP_H + k]);
    }
}

__global__ void k_mlp_batch_idx(int *idx, int nbatch, long long n, uint64_t seed, int step) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbatch) return;
    Philox r;
    philox_init(r, seed, 3, (uint64_t)step, (uint32_t)i);
    philox_refill(r);
    uint64_t u = ((uint64_t)r.b0 << 32) | r.b1;
    idx[i] = (int)(u % (uint64_t)n);
}

__global__ void k_mlp_zero(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

struct dr_edge_mlp {
    float *P = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
    unsigned short *W1T = nullptr, *W2T = nullptr;
    int *idx = nullptr;
    int idx_cap = 0;
    int t = 0;
};

static void mlp_free(dr_edge_mlp *q) {
    void *ps[] = {q->P, q->grad, q->m, q->v, q->loss, q->W1T, q->W2T, q->idx};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete q;
}

extern "C" {

dr_status dr_mlp_create(uint64_t seed, dr_edge_mlp **out) {
    if (!out) return dr_internal_fail(DR_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    dr_edge_mlp *q = new (std::nothrow) dr_edge_mlp();
    if (!q) return dr_internal_fail(DR_ERROR_OUT_OF_MEMORY, "host allocation failed");
    bool ok = hipMalloc(&q->P, P_COUNT * 4) == hipSuccess && hipMalloc(&q->grad, P_COUNT * 4) == hipSuccess &&
              hipMalloc(&q->m, P_COUNT * 4) == hipSuccess && hipMalloc(&q->v, P_COUNT * 4) == hipSuccess &&
              hipMalloc(&q->loss, 4) == hipSuccess && hipMalloc(&q->W1T, MLP_H * MLP_INP * 2) == hipSuccess &&
              hipMalloc(&q->W2T, MLP_H * MLP_H * 2) == hipSuccess;
    if (!ok) {
        mlp_free(q);
        return dr_internal_fail(DR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    }
    // He-normal weights (Box-Muller on a 64-bit LCG), zero biases; the output
    // layer starts small so that the first proposal is nearly uniform
    std::vector<float> h(P_COUNT, 0.f);
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int i = 0; i < MLP_IN * MLP_H; i++) h[P_W1 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
    for (int i = 0; i < MLP_H * MLP_H; i++) h[P_W2 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
    for (int i = 0; i < MLP_H; i++) h[P_W3 + i] = (float)(normal() * 0.01);
    std::vector<float> z(P_COUNT, 0.f);
    ok = hipMemcpy(q->P, h.data(), P_COUNT * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(q->grad, z.data(), P_COUNT * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(q->m, z.data(), P_COUNT * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(q->v, z.data(), P_COUNT * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        mlp_free(q);
        return dr_internal_fail(DR_ERROR_DEVICE, "parameter upload failed");
    }
    k_mlp_pack<<<(MLP_H * MLP_H + 255) / 256, 256>>>(q->P, q->W1T, q->W2T);
    if (hipDeviceSynchronize() != hipSuccess) {
        mlp_free(q);
        return dr_internal_fail(DR_ERROR_DEVICE, "weight packing failed");
    }
    *out = q;
    return DR_OK;
}

void dr_mlp_destroy(dr_edge_mlp *q) {
    if (q) mlp_free(q);
}

dr_status dr_mlp_params(dr_edge_mlp *q, float *get, const float *set) {
    if (!q) return dr_internal_fail(DR_ERROR_INVALID_ARGUMENT, "null network");
    if (set) {
        if (hipMemcpy(q->P, set, P_COUNT * 4, hipMemcpyHostToDevice) != hipSuccess)
            return dr_internal_fail(DR_ERROR_DEVICE, "parameter upload failed");
        k_mlp_pack<<<(MLP_H * MLP_H + 255) / 256, 256>>>(q->P, q->W1T, q->W2T);
    }
    if (get && hipMemcpy(get, q->P, P_COUNT * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return dr_internal_fail(DR_ERROR_DEVICE, "parameter download failed");
    if (hipDeviceSynchronize() != hipSuccess) return dr_internal_fail(DR_ERROR_DEVICE, "synchronize failed");
    return DR_OK;
}

dr_status dr_mlp_infer(dr_edge_mlp *q, const float *features, int64_t n, float *logits, void *stream) {
    if (!q || n < 0 || (n > 0 && (!features || !logits)))
        return dr_internal_fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return DR_OK;
    long long waves = (n + 15) / 16;
    k_mlp_infer<<<(unsigned)((waves + 3) / 4), 256, 0, (hipStream_t)stream>>>(features, n, q->W1T, q->W2T, q->P,
                                                                              logits);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DR_OK : dr_internal_fail(DR_ERROR_DEVICE, hipGetErrorString(e));
}

dr_status dr_mlp_train(dr_edge_mlp *q, const float *features, const float *targets, const float *weights,
                       int64_t n, int steps, int batch, float lr, uint64_t seed, float *last_loss, void *stream) {
    if (!q || !features || !targets || !weights || n <= 0 || steps < 0 || batch <= 0 || n >= (1ll << 31))
        return dr_internal_fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    hipStream_t hs = (hipStream_t)stream;
    int nb = (batch + MB - 1) / MB * MB;
    if (nb > q->idx_cap) {
        if (q->idx) (void)hipFree(q->idx);
        q->idx = nullptr;
        if (hipMalloc(&q->idx, nb * sizeof(int)) != hipSuccess)
            return dr_internal_fail(DR_ERROR_OUT_OF_MEMORY, "minibatch buffer");
        q->idx_cap = nb;
    }
    for (int s = 0; s < steps; s++) {
        k_mlp_batch_idx<<<(nb + 255) / 256, 256, 0, hs>>>(q->idx, nb, n, seed, q->t);
        k_mlp_zero<<<1, 64, 0, hs>>>(q->loss, 1);
        k_mlp_grad<<<nb / MB, 256, 0, hs>>>(features, targets, weights, q->idx, nb, q->P, q->grad, q->loss,
                                            1.f / (float)nb);
        q->t++;
        k_mlp_adam<<<(P_COUNT + 255) / 256, 256, 0, hs>>>(q->P, q->grad, q->m, q->v, q->t, lr, q->W1T, q->W2T);
    }
    k_mlp_pack<<<(MLP_H * MLP_H + 255) / 256, 256, 0, hs>>>(q->P, q->W1T, q->W2T);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return dr_internal_fail(DR_ERROR_DEVICE, hipGetErrorString(e));
    if (last_loss) {
        if (hipMemcpyAsync(last_loss, q->loss, 4, hipMemcpyDeviceToHost, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return dr_internal_fail(DR_ERROR_DEVICE, "loss read failed");
    }
    return DR_OK;
}

}  // extern "C"

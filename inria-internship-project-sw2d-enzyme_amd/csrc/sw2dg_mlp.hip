// sw2dg_mlp.hip -- the learned warm start: a small MLP (dense layers, GELU
// between them) trained and evaluated with hand-written BF16 MFMA kernels
// (FP32 accumulation) on gfx950.  It maps observed gauge series to an
// estimate of the calibration controls; the 4D-Var refines from there.
//
// GEMM: C[M x N] = op(A)[M x K] op(B)[K x N], A/B in BF16 (row-major, op = T
// reads the transpose), 64 x 64 output tile per workgroup of 4 waves, each
// wave a 32 x 32 block of v_mfma_f32_32x32x16_bf16, K staged through LDS 32 at
// a time.  Epilogues fuse the bias, GELU (forward), GELU' (backward) and the
// BF16 re-quantisation of activations.  Training: mean squared error, Adam
// with FP32 master weights and BF16 working copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16_t;

__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_BACK_ACT = 2, EPI_PLAIN = 3 };

struct Epi {
    const float *bias;   // [N] (HIDDEN, OUT)
    float *z;            // fp32 result [M x N] (HIDDEN: pre-activation; OUT: output; BACK_ACT: dZ; PLAIN: C)
    bf16_t *h;           // bf16 copy [M x N] (HIDDEN: GELU(z); BACK_ACT: dZ)
    const float *zsave;  // BACK_ACT: the layer's saved pre-activation [M x N]
    int ldz;
};

constexpr int TM = 64, TN = 64, TK = 32, PADK = 8;

template <bool TA, bool TB, int EPIK>
__global__ __launch_bounds__(256) void k_gemm(const bf16_t *__restrict__ A, const bf16_t *__restrict__ B,
                                             int M, int N, int K, int lda, int ldb, Epi ep) {
    __shared__ __attribute__((aligned(16))) bf16_t sA[TM][TK + PADK];  // [m][k]
    __shared__ __attribute__((aligned(16))) bf16_t sB[TN][TK + PADK];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int mb = blockIdx.y * TM, nb = blockIdx.x * TN;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int kb = 0; kb < K; kb += TK) {
        // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread
        for (int q = tid; q < TM * TK; q += 256) {
            int m, k;
            if (TA) { m = q % TM; k = q / TM; }  // walk m fastest: contiguous in A^T storage
            else { k = q % TK; m = q / TK; }
            const int gm = mb + m, gk = kb + k;
            bf16_t v = 0;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            sA[m][k] = v;
        }
        for (int q = tid; q < TN * TK; q += 256) {
            int n, k;
            if (TB) { k = q % TK; n = q / TK; }
            else { n = q % TN; k = q / TN; }
            const int gn = nb + n, gk = kb + k;
            bf16_t v = 0;
            if (gn < N && gk < K) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            sB[n][k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            const int r = lane & 31, k0 = kk + 8 * (lane >> 5);
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sA[32 * wm + r][k0]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sB[32 * wn + r][k0]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: lane holds column (lane % 32), rows 8 (v / 4) + 4 (lane / 32) + v % 4
    const int col = nb + 32 * wn + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = mb + 32 * wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row >= M || col >= N) continue;
        const size_t o = (size_t)row * ep.ldz + col;
        const float x = acc[v];
        if (EPIK == EPI_HIDDEN) {
            const float z = x + ep.bias[col];
            if (ep.z) ep.z[o] = z;  // kept for the backward pass only
            ep.h[o] = f2bf(gelu(z));
        } else if (EPIK == EPI_OUT) {
            ep.z[o] = x + ep.bias[col];
        } else if (EPIK == EPI_BACK_ACT) {
            const float dz = x * gelu_d(ep.zsave[o]);
            ep.z[o] = dz;
            ep.h[o] = f2bf(dz);
        } else {
            ep.z[o] = x;
        }
    }
}

__global__ void k_to_bf16(const float *__restrict__ x, bf16_t *__restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// column sums (bias gradient), fixed order over the rows
__global__ void k_colsum(const float *__restrict__ x, int M, int N, float *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s += x[(size_t)m * N + n];
    out[n] = s;
}

// mean squared error over the batch: dY = 2 (Y - T) / (M N); per-row partial losses
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, int M, int N,
                      float *__restrict__ dy, bf16_t *__restrict__ dyb, float *__restrict__ row_loss) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float sc = 2.0f / ((float)M * (float)N);
    float l = 0.0f;
    for (int n = 0; n < N; ++n) {
        const size_t o = (size_t)m * N + n;
        const float d = y[o] - t[o];
        l += d * d;
        dy[o] = sc * d;
        dyb[o] = f2bf(sc * d);
    }
    row_loss[m] = l;
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, bf16_t *__restrict__ wb, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// Scratch of sw2dg_mlp_forward_noalloc: the BF16 input and the hidden
// activations of one batch, in a fixed device array (one copy per GPU, held
// until the process ends) so that the call allocates nothing.  Host code uses
// the per-device address from hipGetSymbolAddress, never the array's name.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices] = {};

// Address of the scratch array on `device`, looked up once per device with
// that device current (called from sw2dg_mlp_create).
int scratch_setup(int device, std::string &err) {
    if (device < 0 || device >= kMaxDevices) {
        err = "scratch: device index " + std::to_string(device) + " out of range";
        return SW2DG_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return SW2DG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return SW2DG_ERR_HIP;
    }
    g_scratch_dev[device] = (unsigned char *)p;
    return SW2DG_OK;
}

unsigned char *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct sw2dg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int L = 0, max_batch = 0;
    std::vector<int> dims;
    // per layer: W [in x out] fp32 master + bf16 copy, b [out]; Adam moments
    std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<bf16_t *> Wb;
    // activations: X bf16 [B x d0]; per layer: Z fp32, H bf16 (hidden), dZ fp32 / bf16
    bf16_t *xb = nullptr;
    std::vector<float *> Z, dZ;
    std::vector<bf16_t *> H, dZb;
    float *X32 = nullptr, *T32 = nullptr, *row_loss = nullptr;
    int64_t step = 0;
    std::vector<void *> allocs;
    template <class T_>
    T_ *dalloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T_)) != hipSuccess) return nullptr;
        hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T_));
        allocs.push_back(p);
        return (T_ *)p;
    }
    ~sw2dg_mlp() {
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
int mfail(sw2dg_mlp *m, int code, const char *msg) {
    if (m) m->err = msg;
    return code;
}
#define MHIP(m, x)                                                                      \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) return mfail(m, SW2DG_ERR_HIP, hipGetErrorString(e_));   \
    } while (0)

template <bool TA, bool TB, int EPIK>
void gemm_on(hipStream_t s, const bf16_t *A, const bf16_t *B, int M, int N, int K, int lda, int ldb,
             const Epi &ep) {
    const dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPIK>), grid, dim3(256), 0, s, A, B, M, N, K, lda, ldb, ep);
}

template <bool TA, bool TB, int EPIK>
void gemm(sw2dg_mlp *m, const bf16_t *A, const bf16_t *B, int M, int N, int K, int lda, int ldb,
          const Epi &ep) {
    gemm_on<TA, TB, EPIK>(m->stream, A, B, M, N, K, lda, ldb, ep);
}

// bytes of scratch a forward of `batch` rows takes: the BF16 input, then the
// BF16 activations of each hidden layer (the output goes to the caller)
size_t noalloc_bytes(const sw2dg_mlp *m, size_t batch) {
    size_t b = align256(batch * m->dims[0] * sizeof(bf16_t));
    for (int l = 0; l + 1 < m->L; ++l) b += align256(batch * m->dims[l + 1] * sizeof(bf16_t));
    return b;
}

size_t n_params(const sw2dg_mlp *m) {
    size_t n = 0;
    for (int l = 0; l < m->L; ++l) n += (size_t)m->dims[l] * m->dims[l + 1] + m->dims[l + 1];
    return n;
}

// forward of a batch whose inputs are in m->xb; output fp32 in m->Z[L-1]
void forward(sw2dg_mlp *m, int B) {
    const bf16_t *in = m->xb;
    for (int l = 0; l < m->L; ++l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        Epi ep{m->b[l], m->Z[l], l + 1 < m->L ? m->H[l] : nullptr, nullptr, N};
        if (l + 1 < m->L) gemm<false, false, EPI_HIDDEN>(m, in, m->Wb[l], B, N, K, K, N, ep);
        else gemm<false, false, EPI_OUT>(m, in, m->Wb[l], B, N, K, K, N, ep);
        in = m->H[l];
    }
}

// backward from dZ[L-1] (fp32 + bf16 already set): weight / bias gradients
void backward(sw2dg_mlp *m, int B) {
    for (int l = m->L - 1; l >= 0; --l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const bf16_t *in = l ? m->H[l - 1] : m->xb;
        // dW = in^T dZ  [K x N]
        Epi ew{nullptr, m->gW[l], nullptr, nullptr, N};
        gemm<true, false, EPI_PLAIN>(m, in, m->dZb[l], K, N, B, K, N, ew);
        hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, m->stream, m->dZ[l], B, N, m->gb[l]);
        if (l > 0) {  // dZ_{l-1} = (dZ W^T) * GELU'(Z_{l-1})
            Epi ed{nullptr, m->dZ[l - 1], m->dZb[l - 1], m->Z[l - 1], K};
            gemm<false, true, EPI_BACK_ACT>(m, m->dZb[l], m->Wb[l], B, K, N, N, N, ed);
        }
    }
}

int load_batch(sw2dg_mlp *m, int B, const float *x, const float *t) {
    const int d0 = m->dims[0], dl = m->dims[m->L];
    MHIP(m, hipMemcpyAsync(m->X32, x, (size_t)B * d0 * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const size_t n = (size_t)B * d0;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, m->X32, m->xb, n);
    if (t) MHIP(m, hipMemcpyAsync(m->T32, t, (size_t)B * dl * sizeof(float), hipMemcpyHostToDevice, m->stream));
    return SW2DG_OK;
}

int loss_and_backward(sw2dg_mlp *m, int B, float *loss) {
    const int dl = m->dims[m->L];
    hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->Z[m->L - 1], m->T32, B, dl,
                       m->dZ[m->L - 1], m->dZb[m->L - 1], m->row_loss);
    backward(m, B);
    if (loss) {
        std::vector<float> rl(B);
        MHIP(m, hipMemcpyAsync(rl.data(), m->row_loss, B * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (int i = 0; i < B; ++i) s += rl[i];
        *loss = (float)(s / ((double)B * dl));
    }
    return SW2DG_OK;
}
}  // namespace

extern "C" {

int sw2dg_mlp_create(int32_t n_layers, const int32_t *dims, int32_t max_batch, int32_t device,
                     sw2dg_mlp **out) {
    if (!out) return SW2DG_ERR_ARG;
    *out = nullptr;
    if (n_layers < 1 || n_layers > 16 || !dims || max_batch < 1) return SW2DG_ERR_ARG;
    for (int l = 0; l <= n_layers; ++l)
        if (dims[l] < 1 || dims[l] > 65536) return SW2DG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SW2DG_ERR_NO_DEVICE;
    auto *m = new sw2dg_mlp();
    if (device < 0) hipGetDevice(&m->device);
    else m->device = device;
    if (hipSetDevice(m->device) != hipSuccess ||
        hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return SW2DG_ERR_HIP;
    }
    if (scratch_setup(m->device, m->err) != SW2DG_OK) {
        delete m;
        return SW2DG_ERR_HIP;
    }
    m->L = n_layers;
    m->max_batch = max_batch;
    m->dims.assign(dims, dims + n_layers + 1);
    for (int l = 0; l < n_layers; ++l) {
        const size_t nw = (size_t)dims[l] * dims[l + 1], nb = dims[l + 1];
        m->W.push_back(m->dalloc<float>(nw));
        m->Wb.push_back(m->dalloc<bf16_t>(nw));
        m->b.push_back(m->dalloc<float>(nb));
        m->gW.push_back(m->dalloc<float>(nw));
        m->gb.push_back(m->dalloc<float>(nb));
        m->mW.push_back(m->dalloc<float>(nw));
        m->vW.push_back(m->dalloc<float>(nw));
        m->mb.push_back(m->dalloc<float>(nb));
        m->vb.push_back(m->dalloc<float>(nb));
        m->Z.push_back(m->dalloc<float>((size_t)max_batch * dims[l + 1]));
        m->dZ.push_back(m->dalloc<float>((size_t)max_batch * dims[l + 1]));
        m->H.push_back(m->dalloc<bf16_t>((size_t)max_batch * dims[l + 1]));
        m->dZb.push_back(m->dalloc<bf16_t>((size_t)max_batch * dims[l + 1]));
    }
    m->xb = m->dalloc<bf16_t>((size_t)max_batch * dims[0]);
    m->X32 = m->dalloc<float>((size_t)max_batch * dims[0]);
    m->T32 = m->dalloc<float>((size_t)max_batch * dims[n_layers]);
    m->row_loss = m->dalloc<float>(max_batch);
    for (void *p : m->allocs)
        if (!p) {
            delete m;
            return SW2DG_ERR_HIP;
        }
    *out = m;
    return SW2DG_OK;
}

int64_t sw2dg_mlp_n_params(sw2dg_mlp *m) { return m ? (int64_t)n_params(m) : 0; }

int sw2dg_mlp_set_params(sw2dg_mlp *m, const float *params) {
    if (!m || !params) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    size_t o = 0;
    for (int l = 0; l < m->L; ++l) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        MHIP(m, hipMemcpy(m->W[l], params + o, nw * sizeof(float), hipMemcpyHostToDevice));
        o += nw;
        MHIP(m, hipMemcpy(m->b[l], params + o, nb * sizeof(float), hipMemcpyHostToDevice));
        o += nb;
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, m->W[l], m->Wb[l], nw);
        MHIP(m, hipMemsetAsync(m->mW[l], 0, nw * sizeof(float), m->stream));
        MHIP(m, hipMemsetAsync(m->vW[l], 0, nw * sizeof(float), m->stream));
        MHIP(m, hipMemsetAsync(m->mb[l], 0, nb * sizeof(float), m->stream));
        MHIP(m, hipMemsetAsync(m->vb[l], 0, nb * sizeof(float), m->stream));
    }
    m->step = 0;
    MHIP(m, hipStreamSynchronize(m->stream));
    return SW2DG_OK;
}

int sw2dg_mlp_get_params(sw2dg_mlp *m, float *params) {
    if (!m || !params) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipStreamSynchronize(m->stream));
    size_t o = 0;
    for (int l = 0; l < m->L; ++l) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        MHIP(m, hipMemcpy(params + o, m->W[l], nw * sizeof(float), hipMemcpyDeviceToHost));
        o += nw;
        MHIP(m, hipMemcpy(params + o, m->b[l], nb * sizeof(float), hipMemcpyDeviceToHost));
        o += nb;
    }
    return SW2DG_OK;
}

int sw2dg_mlp_forward(sw2dg_mlp *m, int32_t batch, const float *x, float *y) {
    if (!m || !x || !y || batch < 1 || batch > m->max_batch) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    int st = load_batch(m, batch, x, nullptr);
    if (st) return st;
    forward(m, batch);
    MHIP(m, hipGetLastError());
    MHIP(m, hipMemcpyAsync(y, m->Z[m->L - 1], (size_t)batch * m->dims[m->L] * sizeof(float),
                           hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return SW2DG_OK;
}

int sw2dg_mlp_forward_noalloc(sw2dg_mlp *m, int32_t batch, const float *x, float *y, void *stream) {
    if (!m || !x || !y || batch < 1) return SW2DG_ERR_ARG;
    const size_t need = noalloc_bytes(m, (size_t)batch);
    if (need > kScratchBytes) {
        const size_t per_row = noalloc_bytes(m, 1);
        m->err = "sw2dg_mlp_forward_noalloc: batch " + std::to_string(batch) + " needs " +
                 std::to_string(need) + " bytes of scratch, above the maximum of " +
                 std::to_string(kScratchBytes) + " (about " + std::to_string(kScratchBytes / per_row) +
                 " rows for these layer widths)";
        return SW2DG_ERR_ARG;
    }
    MHIP(m, hipSetDevice(m->device));
    unsigned char *base = scratch_of(m->device);
    if (!base) return mfail(m, SW2DG_ERR_HIP, "sw2dg_mlp_forward_noalloc: no scratch on this device");
    hipStream_t s = stream ? (hipStream_t)stream : m->stream;
    // every scratch byte read below is written first in the same call: the
    // input conversion fills xb, each hidden GEMM fills its activations
    // before the next layer reads them, so nothing needs clearing
    const size_t n0 = (size_t)batch * m->dims[0];
    bf16_t *xb = (bf16_t *)base;
    size_t off = align256(n0 * sizeof(bf16_t));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s, x, xb, n0);
    const bf16_t *in = xb;
    for (int l = 0; l < m->L; ++l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        if (l + 1 < m->L) {
            bf16_t *h = (bf16_t *)(base + off);
            off += align256((size_t)batch * N * sizeof(bf16_t));
            Epi ep{m->b[l], nullptr, h, nullptr, N};
            gemm_on<false, false, EPI_HIDDEN>(s, in, m->Wb[l], batch, N, K, K, N, ep);
            in = h;
        } else {
            Epi ep{m->b[l], y, nullptr, nullptr, N};
            gemm_on<false, false, EPI_OUT>(s, in, m->Wb[l], batch, N, K, K, N, ep);
        }
    }
    MHIP(m, hipGetLastError());
    if (!stream) MHIP(m, hipStreamSynchronize(m->stream));
    return SW2DG_OK;
}

int sw2dg_mlp_grad(sw2dg_mlp *m, int32_t batch, const float *x, const float *t, float *loss,
                   float *grads) {
    if (!m || !x || !t || batch < 1 || batch > m->max_batch) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    int st = load_batch(m, batch, x, t);
    if (st) return st;
    forward(m, batch);
    if ((st = loss_and_backward(m, batch, loss))) return st;
    MHIP(m, hipGetLastError());
    MHIP(m, hipStreamSynchronize(m->stream));
    if (grads) {
        size_t o = 0;
        for (int l = 0; l < m->L; ++l) {
            const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
            MHIP(m, hipMemcpy(grads + o, m->gW[l], nw * sizeof(float), hipMemcpyDeviceToHost));
            o += nw;
            MHIP(m, hipMemcpy(grads + o, m->gb[l], nb * sizeof(float), hipMemcpyDeviceToHost));
            o += nb;
        }
    }
    return SW2DG_OK;
}

int sw2dg_mlp_train_step(sw2dg_mlp *m, int32_t batch, const float *x, const float *t, float lr,
                         float *loss) {
    if (!m || !x || !t || batch < 1 || batch > m->max_batch || !(lr > 0.f)) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    int st = load_batch(m, batch, x, t);
    if (st) return st;
    forward(m, batch);
    if ((st = loss_and_backward(m, batch, loss))) return st;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    for (int l = 0; l < m->L; ++l) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, m->W[l], m->gW[l],
                           m->mW[l], m->vW[l], m->Wb[l], nw, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, m->stream, m->b[l], m->gb[l],
                           m->mb[l], m->vb[l], (bf16_t *)nullptr, nb, lr, b1, b2, eps, c1, c2);
    }
    MHIP(m, hipGetLastError());
    if (!loss) MHIP(m, hipStreamSynchronize(m->stream));
    return SW2DG_OK;
}

const char *sw2dg_mlp_last_error(sw2dg_mlp *m) { return m ? m->err.c_str() : "null mlp"; }
void sw2dg_mlp_destroy(sw2dg_mlp *m) { delete m; }

}  // extern "C"

// sw2dg_mlp.hip -- the learned warm start: a small MLP (dense layers, GELU
// between them) trained and evaluated with hand-written BF16 MFMA kernels
// (FP32 accumulation) on gfx950.  It maps observed gauge series to an
// estimate of the calibration controls; the 4D-Var refines from there.
//
// GEMM: C[M x N] = op(A)[M x K] op(B)[K x N], A/B in BF16 (row-major, op = T
// reads the transpose), a 64 x 64 or 128 x 128 output tile per workgroup of 4
// waves on v_mfma_f32_32x32x16_bf16, K staged through LDS 64 at a time with
// 16-byte loads and a register prefetch; small grids split K (fixed-order
// sum of the splits).  Epilogues fuse the bias, GELU (forward), GELU' (backward) and the
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

constexpr int TK = 64, LDK = TK + 8;  // LDS rows of 72 BF16 (144 bytes, 16-byte aligned)
constexpr int TM = 64, TN = 64;       // the small output tile (split-K allowed)
// Split-K: tiles x splits never exceed this, so the FP32 partials of one GEMM
// hold at most kMaxSplitTiles x 64 x 64 floats (8 MiB).
constexpr int kMaxSplitTiles = 512;

// 8 consecutive BF16 from global memory, zero past the first `cnt`; one
// 16-byte load when all 8 are in range and the address is 16-byte aligned
__device__ __forceinline__ uint4 ld8(const bf16_t *p, int cnt, bool vec) {
    if (vec && cnt == 8) return *reinterpret_cast<const uint4 *>(p);
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (i < cnt) w[i >> 1] |= (uint32_t)p[i] << (16 * (i & 1));
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ bf16_t half_of(const uint4 &r, int i) {
    const uint32_t w = i < 2 ? r.x : i < 4 ? r.y : i < 6 ? r.z : r.w;
    return (bf16_t)(w >> (16 * (i & 1)));
}

template <int EPIK>
__device__ __forceinline__ void epi_store(const Epi &ep, int row, int col, float x) {
    const size_t o = (size_t)row * ep.ldz + col;
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

// One BT x BT output tile per workgroup (BT = 64 or 128; 4 waves, each a
// BT/2 x BT/2 block of 32 x 32 MFMA tiles), over the k-tiles
// [blockIdx.z * kt_per, +kt_per) of 64.  Each thread stages BT/32
// 8-element vectors of A and of B per k-tile (16-byte loads along the
// contiguous index; a vector along m or n is written to LDS transposed), and
// loads the next k-tile into registers before the MFMAs of the current one.
// Every output element accumulates its k in the same order whatever BT is.
// part == nullptr: the epilogue writes the result; else the raw FP32 sums of
// this split go to part[blockIdx.z][M][N] and k_split_sum finishes.
template <bool TA, bool TB, int EPIK, int BT>
__global__ __launch_bounds__(256) void k_gemm(const bf16_t *__restrict__ A, const bf16_t *__restrict__ B,
                                             int M, int N, int K, int lda, int ldb, int kt_per,
                                             float *__restrict__ part, Epi ep) {
    constexpr int NV = BT / 32, WT = BT / 2, NB = WT / 32;
    __shared__ __attribute__((aligned(16))) bf16_t sA[BT][LDK];  // [m][k]
    __shared__ __attribute__((aligned(16))) bf16_t sB[BT][LDK];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int mb = blockIdx.y * BT, nb = blockIdx.x * BT;
    const int nkt = (K + TK - 1) / TK;
    const int kt0 = blockIdx.z * kt_per, kt1 = min(nkt, kt0 + kt_per);
    const bool va = ((lda & 7) == 0) && (((uintptr_t)A & 15) == 0);
    const bool vb = ((ldb & 7) == 0) && (((uintptr_t)B & 15) == 0);
    // vector v of a tile: along k (row r = v / 8 of BT, k = 8 (v % 8)), or
    // along m / n (k-row v / (BT / 8) of TK, m or n = 8 (v % (BT / 8)))
    uint4 ra[NV], rb[NV];
    auto fetch = [&](int kt) {
        const int kb = kt * TK;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int v = tid + 256 * j;
            const int rd = v >> 3, qd = (v & 7) * 8, rt = v / (BT / 8), qt = (v % (BT / 8)) * 8;
            if (!TA) {  // row m = rd of A, k from qd
                const int gm = mb + rd, gk = kb + qd;
                const int cnt = gm < M ? max(0, min(8, K - gk)) : 0;
                ra[j] = ld8(A + (size_t)gm * lda + gk, cnt, va);
            } else {    // row k = rt of A^T storage, m from qt
                const int gk = kb + rt, gm = mb + qt;
                const int cnt = gk < K ? max(0, min(8, M - gm)) : 0;
                ra[j] = ld8(A + (size_t)gk * lda + gm, cnt, va);
            }
            if (TB) {   // row n = rd of B^T storage, k from qd
                const int gn = nb + rd, gk = kb + qd;
                const int cnt = gn < N ? max(0, min(8, K - gk)) : 0;
                rb[j] = ld8(B + (size_t)gn * ldb + gk, cnt, vb);
            } else {    // row k = rt of B, n from qt
                const int gk = kb + rt, gn = nb + qt;
                const int cnt = gk < K ? max(0, min(8, N - gn)) : 0;
                rb[j] = ld8(B + (size_t)gk * ldb + gn, cnt, vb);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int v = tid + 256 * j;
            const int rd = v >> 3, qd = (v & 7) * 8, rt = v / (BT / 8), qt = (v % (BT / 8)) * 8;
            if (!TA) {
                *reinterpret_cast<uint4 *>(&sA[rd][qd]) = ra[j];
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) sA[qt + i][rt] = half_of(ra[j], i);
            }
            if (TB) {
                *reinterpret_cast<uint4 *>(&sB[rd][qd]) = rb[j];
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) sB[qt + i][rt] = half_of(rb[j], i);
            }
        }
    };
    f32x16 acc[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    if (kt0 < kt1) fetch(kt0);
    for (int kt = kt0; kt < kt1; ++kt) {
        stash();
        __syncthreads();
        if (kt + 1 < kt1) fetch(kt + 1);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            const int r = lane & 31, k0 = kk + 8 * (lane >> 5);
            bf16x8 a[NB], b[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(&sA[WT * wm + 32 * i + r][k0]);
#pragma unroll
            for (int j = 0; j < NB; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(&sB[WT * wn + 32 * j + r][k0]);
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane holds column (lane % 32), rows 8 (v / 4) + 4 (lane / 32) + v % 4 of each 32 x 32 tile
    float *pz = part ? part + (size_t)blockIdx.z * M * N : nullptr;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int col = nb + WT * wn + 32 * j + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mb + WT * wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row >= M || col >= N) continue;
                if (pz) pz[(size_t)row * N + col] = acc[i][j][v];
                else epi_store<EPIK>(ep, row, col, acc[i][j][v]);
            }
        }
}

// the splits of a split-K GEMM summed in order (split 0 first), then the epilogue
template <int EPIK>
__global__ __launch_bounds__(256) void k_split_sum(const float *__restrict__ part, int S, int M, int N, Epi ep) {
    const size_t MN = (size_t)M * N, i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= MN) return;
    float x = part[i];
    for (int s = 1; s < S; ++s) x += part[(size_t)s * MN + i];
    epi_store<EPIK>(ep, (int)(i / N), (int)(i % N), x);
}

// training batch: rows idx[0..B) of the resident data set (BF16 inputs, FP32 targets)
__global__ void k_gather_rows(const bf16_t *__restrict__ dx, const float *__restrict__ dt,
                              const int32_t *__restrict__ idx, int B, int d0, int dl,
                              bf16_t *__restrict__ xb, float *__restrict__ tb) {
    const int r = blockIdx.x;
    if (r >= B) return;
    const size_t src = (size_t)idx[r];
    for (int c = threadIdx.x; c < d0; c += blockDim.x) xb[(size_t)r * d0 + c] = dx[src * d0 + c];
    for (int c = threadIdx.x; c < dl; c += blockDim.x) tb[(size_t)r * dl + c] = dt[src * dl + c];
}

__global__ void k_to_bf16(const float *__restrict__ x, bf16_t *__restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// column sums (bias gradient): a workgroup per 64 columns, its 4 waves each
// summing every 4th row in row order (8 loads in flight), the 4 partial sums
// added in wave order -- a fixed order, so the result is deterministic
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ x, int M, int N,
                                               float *__restrict__ out) {
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), w = threadIdx.x >> 6;
    float s = 0.0f;
    if (c < N) {
        int m = w;
        for (; m + 28 < M; m += 32) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = x[(size_t)(m + 4 * j) * N + c];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        for (; m < M; m += 4) s += x[(size_t)m * N + c];
    }
    part[w][threadIdx.x & 63] = s;
    __syncthreads();
    if (w == 0 && c < N) out[c] = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                                  part[3][threadIdx.x];
}

// mean squared error over the batch: dY = 2 (Y - T) / (M N); per-row
// partial losses.  One wave per row: lane j takes columns j, j + 64, ...;
// the lanes' sums meet in a fixed butterfly order.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y, const float *__restrict__ t, int M,
                                            int N, float *__restrict__ dy, bf16_t *__restrict__ dyb,
                                            float *__restrict__ row_loss) {
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (m >= M) return;  // a whole wave leaves together
    const float sc = 2.0f / ((float)M * (float)N);
    float l = 0.0f;
    for (int n = lane; n < N; n += 64) {
        const size_t o = (size_t)m * N + n;
        const float d = y[o] - t[o];
        l += d * d;
        dy[o] = sc * d;
        dyb[o] = f2bf(sc * d);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (lane == 0) row_loss[m] = l;
}

// the weight segments of the parameter array: layer l's W [K x N] at off[l]
struct WSegs {
    int L;
    size_t off[16];
    int K[16], N[16];
};

// Adam over the whole parameter array; the BF16 copies follow: wb in the
// same layout, wbt with every weight segment transposed ([N x K], the
// forward GEMM's k-contiguous operand)
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, bf16_t *__restrict__ wb, bf16_t *__restrict__ wbt, WSegs sg,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = wi;
    const bf16_t q = f2bf(wi);
    wb[i] = q;
    for (int l = 0; l < sg.L; ++l) {
        const size_t e = i - sg.off[l];
        if (i >= sg.off[l] && e < (size_t)sg.K[l] * sg.N[l]) {
            wbt[sg.off[l] + (e % sg.N[l]) * sg.K[l] + e / sg.N[l]] = q;
            break;
        }
    }
}

// W [K x N] FP32 -> its BF16 transpose [N x K]
__global__ void k_to_bf16_t(const float *__restrict__ w, int K, int N, bf16_t *__restrict__ wt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    wt[(i % N) * K + i / N] = f2bf(w[i]);
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
    // FP32 parameters P (per layer W [in x out] then b [out]), gradients G,
    // Adam moments M1 / M2, BF16 copy PB, all with the same padded layout of
    // n_pad elements; per-layer pointers into them
    float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
    bf16_t *PB = nullptr, *PBT = nullptr;  // PBT: the weight segments transposed ([out x in])
    size_t n_pad = 0;
    WSegs segs{};
    std::vector<float *> W, b, gW, gb;
    std::vector<bf16_t *> Wb, WbT;
    // activations: X bf16 [B x d0]; per layer: Z fp32, H bf16 (hidden), dZ fp32 / bf16
    bf16_t *xb = nullptr;
    std::vector<float *> Z, dZ;
    std::vector<bf16_t *> H, dZb;
    float *X32 = nullptr, *T32 = nullptr, *row_loss = nullptr;
    float *part = nullptr;  // FP32 partials of the split-K training GEMMs (kMaxSplitTiles x 64 x 64)
    int64_t step = 0;
    // resident training set (sw2dg_mlp_set_data) and the row indices / row
    // losses of one sw2dg_mlp_train_epoch call; freed on replacement
    bf16_t *data_x = nullptr;
    float *data_t = nullptr;
    int64_t data_n = 0;
    int32_t *ep_rows = nullptr;
    float *ep_loss = nullptr;
    size_t ep_cap = 0;
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
        for (void *p : {(void *)data_x, (void *)data_t, (void *)ep_rows, (void *)ep_loss})
            if (p) (void)hipFree(p);
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

// Splits of the K loop for a GEMM of `tiles` output tiles and `nkt` k-tiles:
// none when the tiles fill the chip (256 CUs); else as many as keep at least
// 2 k-tiles per split and tiles x splits <= kMaxSplitTiles, with no empty split.
int k_splits(int tiles, int nkt, int &kt_per) {
    int S = 1;
    if (tiles < 256 && nkt >= 4) S = std::max(1, std::min(kMaxSplitTiles / tiles, nkt / 2));
    kt_per = (nkt + S - 1) / S;
    return (nkt + kt_per - 1) / kt_per;
}

// C = op(A) op(B) with the epilogue; part != nullptr allows split-K through
// it (training), nullptr keeps one split (inference: every output row is then
// the same whatever the batch it is computed in)
template <bool TA, bool TB, int EPIK>
void gemm_on(hipStream_t s, const bf16_t *A, const bf16_t *B, int M, int N, int K, int lda, int ldb,
             const Epi &ep, float *part) {
    const int tiles_n = (N + TN - 1) / TN, tiles_m = (M + TM - 1) / TM, nkt = (K + TK - 1) / TK;
    if (tiles_n * tiles_m >= 1024) {  // enough 128 x 128 tiles to fill the chip twice over
        hipLaunchKernelGGL((k_gemm<TA, TB, EPIK, 128>), dim3((N + 127) / 128, (M + 127) / 128, 1), dim3(256), 0, s,
                           A, B, M, N, K, lda, ldb, nkt, (float *)nullptr, ep);
        return;
    }
    int kt_per = nkt;
    const int S = part ? k_splits(tiles_n * tiles_m, nkt, kt_per) : 1;
    hipLaunchKernelGGL((k_gemm<TA, TB, EPIK, 64>), dim3(tiles_n, tiles_m, S), dim3(256), 0, s, A, B, M, N, K, lda,
                       ldb, kt_per, S > 1 ? part : nullptr, ep);
    if (S > 1) {
        const size_t MN = (size_t)M * N;
        hipLaunchKernelGGL((k_split_sum<EPIK>), dim3((unsigned)((MN + 255) / 256)), dim3(256), 0, s, part, S, M, N,
                           ep);
    }
}

template <bool TA, bool TB, int EPIK>
void gemm(sw2dg_mlp *m, const bf16_t *A, const bf16_t *B, int M, int N, int K, int lda, int ldb,
          const Epi &ep, bool split) {
    gemm_on<TA, TB, EPIK>(m->stream, A, B, M, N, K, lda, ldb, ep, split ? m->part : nullptr);
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
// (split: the training path, split-K allowed)
void forward(sw2dg_mlp *m, int B, bool split) {
    const bf16_t *in = m->xb;
    for (int l = 0; l < m->L; ++l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        Epi ep{m->b[l], m->Z[l], l + 1 < m->L ? m->H[l] : nullptr, nullptr, N};
        if (l + 1 < m->L) gemm<false, true, EPI_HIDDEN>(m, in, m->WbT[l], B, N, K, K, K, ep, split);
        else gemm<false, true, EPI_OUT>(m, in, m->WbT[l], B, N, K, K, K, ep, split);
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
        gemm<true, false, EPI_PLAIN>(m, in, m->dZb[l], K, N, B, K, N, ew, true);
        hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(256), 0, m->stream, m->dZ[l], B, N, m->gb[l]);
        if (l > 0) {  // dZ_{l-1} = (dZ W^T) * GELU'(Z_{l-1})
            Epi ed{nullptr, m->dZ[l - 1], m->dZb[l - 1], m->Z[l - 1], K};
            gemm<false, true, EPI_BACK_ACT>(m, m->dZb[l], m->Wb[l], B, K, N, N, N, ed, true);
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

// the loss of a batch from its per-row squared errors: summed in row order in double
float batch_loss(const float *rl, int B, int dl) {
    double s = 0.0;
    for (int i = 0; i < B; ++i) s += rl[i];
    return (float)(s / ((double)B * dl));
}

// MSE of the forward's output against m->T32 (per-row sums to row_loss), then the backward
void mse_and_backward(sw2dg_mlp *m, int B, float *row_loss) {
    const int dl = m->dims[m->L];
    hipLaunchKernelGGL(k_mse, dim3((B + 3) / 4), dim3(256), 0, m->stream, m->Z[m->L - 1], m->T32, B, dl,
                       m->dZ[m->L - 1], m->dZb[m->L - 1], row_loss);
    backward(m, B);
}

int loss_and_backward(sw2dg_mlp *m, int B, float *loss) {
    mse_and_backward(m, B, m->row_loss);
    if (loss) {
        std::vector<float> rl(B);
        MHIP(m, hipMemcpyAsync(rl.data(), m->row_loss, B * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        *loss = batch_loss(rl.data(), B, m->dims[m->L]);
    }
    return SW2DG_OK;
}

// one Adam step over every layer's weights and biases (the BF16 weight copies follow)
void adam_step(sw2dg_mlp *m, float lr) {
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    // padding elements have zero gradient and moments: they stay zero
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->n_pad + 255) / 256)), dim3(256), 0, m->stream, m->P, m->G, m->M1,
                       m->M2, m->PB, m->PBT, m->segs, m->n_pad, lr, b1, b2, eps, c1, c2);
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
    m->segs.L = n_layers;
    m->max_batch = max_batch;
    m->dims.assign(dims, dims + n_layers + 1);
    // parameters, gradients and Adam moments each in one array (every
    // layer's W then b, segments padded to 64 elements so that each starts
    // 256-byte aligned), so that one Adam launch covers the whole network
    std::vector<size_t> offW, offb;
    size_t P = 0;
    for (int l = 0; l < n_layers; ++l) {
        offW.push_back(P);
        P += ((size_t)dims[l] * dims[l + 1] + 63) / 64 * 64;
        offb.push_back(P);
        P += ((size_t)dims[l + 1] + 63) / 64 * 64;
    }
    m->n_pad = P;
    m->P = m->dalloc<float>(P);
    m->G = m->dalloc<float>(P);
    m->M1 = m->dalloc<float>(P);
    m->M2 = m->dalloc<float>(P);
    m->PB = m->dalloc<bf16_t>(P);
    m->PBT = m->dalloc<bf16_t>(P);
    if (!m->P || !m->G || !m->M1 || !m->M2 || !m->PB || !m->PBT) {
        delete m;
        return SW2DG_ERR_HIP;
    }
    for (int l = 0; l < n_layers; ++l) {
        m->W.push_back(m->P + offW[l]);
        m->Wb.push_back(m->PB + offW[l]);
        m->WbT.push_back(m->PBT + offW[l]);
        m->segs.off[l] = offW[l];
        m->segs.K[l] = dims[l];
        m->segs.N[l] = dims[l + 1];
        m->b.push_back(m->P + offb[l]);
        m->gW.push_back(m->G + offW[l]);
        m->gb.push_back(m->G + offb[l]);
        m->Z.push_back(m->dalloc<float>((size_t)max_batch * dims[l + 1]));
        m->dZ.push_back(m->dalloc<float>((size_t)max_batch * dims[l + 1]));
        m->H.push_back(m->dalloc<bf16_t>((size_t)max_batch * dims[l + 1]));
        m->dZb.push_back(m->dalloc<bf16_t>((size_t)max_batch * dims[l + 1]));
    }
    m->xb = m->dalloc<bf16_t>((size_t)max_batch * dims[0]);
    m->X32 = m->dalloc<float>((size_t)max_batch * dims[0]);
    m->T32 = m->dalloc<float>((size_t)max_batch * dims[n_layers]);
    m->row_loss = m->dalloc<float>(max_batch);
    m->part = m->dalloc<float>((size_t)kMaxSplitTiles * TM * TN);
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
        hipLaunchKernelGGL(k_to_bf16_t, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, m->W[l],
                           m->dims[l], m->dims[l + 1], m->WbT[l]);
    }
    MHIP(m, hipMemsetAsync(m->M1, 0, m->n_pad * sizeof(float), m->stream));
    MHIP(m, hipMemsetAsync(m->M2, 0, m->n_pad * sizeof(float), m->stream));
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
    forward(m, batch, false);
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
            gemm_on<false, true, EPI_HIDDEN>(s, in, m->WbT[l], batch, N, K, K, K, ep, nullptr);
            in = h;
        } else {
            Epi ep{m->b[l], y, nullptr, nullptr, N};
            gemm_on<false, true, EPI_OUT>(s, in, m->WbT[l], batch, N, K, K, K, ep, nullptr);
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
    forward(m, batch, true);
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
    forward(m, batch, true);
    if ((st = loss_and_backward(m, batch, loss))) return st;
    adam_step(m, lr);
    MHIP(m, hipGetLastError());
    if (!loss) MHIP(m, hipStreamSynchronize(m->stream));
    return SW2DG_OK;
}

int sw2dg_mlp_set_data(sw2dg_mlp *m, int64_t n, const float *x, const float *t) {
    if (!m || !x || !t || n < 1) return SW2DG_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipStreamSynchronize(m->stream));
    const int d0 = m->dims[0], dl = m->dims[m->L];
    if (m->data_x) (void)hipFree(m->data_x);
    if (m->data_t) (void)hipFree(m->data_t);
    m->data_x = nullptr;
    m->data_t = nullptr;
    m->data_n = 0;
    MHIP(m, hipMalloc((void **)&m->data_x, (size_t)n * d0 * sizeof(bf16_t)));
    MHIP(m, hipMalloc((void **)&m->data_t, (size_t)n * dl * sizeof(float)));
    // inputs through the FP32 staging buffer, max_batch rows at a time, with
    // the same conversion as a batch of sw2dg_mlp_train_step
    for (int64_t r = 0; r < n; r += m->max_batch) {
        const size_t rows = (size_t)std::min<int64_t>(m->max_batch, n - r), cnt = rows * d0;
        MHIP(m, hipMemcpyAsync(m->X32, x + (size_t)r * d0, cnt * sizeof(float), hipMemcpyHostToDevice, m->stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m->stream, m->X32,
                           m->data_x + (size_t)r * d0, cnt);
    }
    MHIP(m, hipMemcpyAsync(m->data_t, t, (size_t)n * dl * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipGetLastError());
    MHIP(m, hipStreamSynchronize(m->stream));
    m->data_n = n;
    return SW2DG_OK;
}

int sw2dg_mlp_train_epoch(sw2dg_mlp *m, int32_t batch, int64_t n_steps, const int32_t *rows, float lr,
                          float *step_loss) {
    if (!m || !rows || batch < 1 || batch > m->max_batch || n_steps < 1 || !(lr > 0.f)) return SW2DG_ERR_ARG;
    if (!m->data_x) return mfail(m, SW2DG_ERR_ARG, "sw2dg_mlp_train_epoch: no training set (sw2dg_mlp_set_data)");
    const size_t nr = (size_t)n_steps * batch;
    for (size_t i = 0; i < nr; ++i)
        if (rows[i] < 0 || rows[i] >= m->data_n)
            return mfail(m, SW2DG_ERR_ARG, "sw2dg_mlp_train_epoch: row index outside the training set");
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipStreamSynchronize(m->stream));
    if (nr > m->ep_cap) {
        if (m->ep_rows) (void)hipFree(m->ep_rows);
        if (m->ep_loss) (void)hipFree(m->ep_loss);
        m->ep_rows = nullptr;
        m->ep_loss = nullptr;
        m->ep_cap = 0;
        MHIP(m, hipMalloc((void **)&m->ep_rows, nr * sizeof(int32_t)));
        MHIP(m, hipMalloc((void **)&m->ep_loss, nr * sizeof(float)));
        m->ep_cap = nr;
    }
    MHIP(m, hipMemcpyAsync(m->ep_rows, rows, nr * sizeof(int32_t), hipMemcpyHostToDevice, m->stream));
    const int d0 = m->dims[0], dl = m->dims[m->L];
    for (int64_t k = 0; k < n_steps; ++k) {
        hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(256), 0, m->stream, m->data_x, m->data_t,
                           m->ep_rows + (size_t)k * batch, batch, d0, dl, m->xb, m->T32);
        forward(m, batch, true);
        mse_and_backward(m, batch, m->ep_loss + (size_t)k * batch);
        adam_step(m, lr);
    }
    MHIP(m, hipGetLastError());
    if (step_loss) {
        std::vector<float> rl(nr);
        MHIP(m, hipMemcpyAsync(rl.data(), m->ep_loss, nr * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        for (int64_t k = 0; k < n_steps; ++k) step_loss[k] = batch_loss(rl.data() + (size_t)k * batch, batch, dl);
    } else {
        MHIP(m, hipStreamSynchronize(m->stream));
    }
    return SW2DG_OK;
}

const char *sw2dg_mlp_last_error(sw2dg_mlp *m) { return m ? m->err.c_str() : "null mlp"; }
void sw2dg_mlp_destroy(sw2dg_mlp *m) { delete m; }

}  // extern "C"

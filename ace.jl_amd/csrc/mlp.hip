// libacemlp.so — the learned surrogate's network: an ensemble of MLPs
// (IN -> 128 -> 128 -> 128 -> 1, SiLU) with BF16 weights and activations and
// FP32 accumulation on the gfx950 MFMA units (v_mfma_f32_16x16x32_bf16).
// Hand-written forward, backward (weight and input gradients), gradient
// reduction and Adam kernels.  No reference counterpart: the reference has no
// learned part (ACE.jl is a linear model).
//
// MFMA 16x16x32 BF16 operand layout (checked by acemlp_layout_probe):
//   A (16 x 32, row-major M x K): lane l holds A[l % 16][8 (l / 16) + 0..7]
//   B (32 x 16, K x N):           lane l holds B[8 (l / 16) + 0..7][l % 16]
//   D (16 x 16 FP32):             lane l holds D[4 (l / 16) + 0..3][l % 16]

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/acemlp.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define INP 160   // padded input width
#define HID 128   // hidden width
#define TROWS 64  // rows per workgroup tile

// parameter offsets of one member (FP32 master copy)
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + HID * INP;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID * HID;
constexpr int OFF_W4 = OFF_B3 + HID;
constexpr int OFF_B4 = OFF_W4 + HID;
constexpr int NPARAM = ((OFF_B4 + 1 + 15) / 16) * 16;
// BF16 copies of one member: W1, W2, W3 row-major [out][in] and transposed [in][out]
constexpr int BOFF_W1 = 0;
constexpr int BOFF_W2 = BOFF_W1 + HID * INP;
constexpr int BOFF_W3 = BOFF_W2 + HID * HID;
constexpr int BOFF_T1 = BOFF_W3 + HID * HID;
constexpr int BOFF_T2 = BOFF_T1 + INP * HID;
constexpr int BOFF_T3 = BOFF_T2 + HID * HID;
constexpr int NBF = BOFF_T3 + HID * HID;

static thread_local std::string g_err;
static int set_err(const std::string& s) {
    g_err = s;
    return -1;
}
#define HIP_TRY(expr)                                                           \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) return set_err(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

// ---------------------------------------------------------------------------
// layout probe: D = A B with A[i][k] = (3i + k) % 7 - 3, B[k][j] = (5k + 2j) % 9 - 4
// (small integers: exact in BF16 and in the FP32 sums)

__global__ void k_layout_probe(float* out) {
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int q = 0; q < 8; ++q) {
        const int i = l % 16, k = 8 * (l / 16) + q, j = l % 16;
        a[q] = (bf16)(float)((3 * i + k) % 7 - 3);
        b[q] = (bf16)(float)((5 * k + 2 * j) % 9 - 4);
    }
    f32x4 d = {0, 0, 0, 0};
    d = mfma(a, b, d);
    for (int r = 0; r < 4; ++r) out[(4 * (l / 16) + r) * 16 + (l % 16)] = d[r];
}

// ---------------------------------------------------------------------------
// forward: grid (ntile, nmemb), block 256 (4 waves, 16 rows each)

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_fwd(int64_t nrows, const bf16* __restrict__ X,
                                             const bf16* __restrict__ wb, const float* __restrict__ params,
                                             float* __restrict__ y, float* __restrict__ Z,
                                             bf16* __restrict__ Hs) {
    __shared__ __attribute__((aligned(16))) bf16 xs[TROWS * INP];
    __shared__ __attribute__((aligned(16))) bf16 hs[2][TROWS * HID];
    const int m = blockIdx.y;
    const int nmemb = gridDim.y;
    const bf16* W = wb + (size_t)m * NBF;
    const float* P = params + (size_t)m * NPARAM;
    const int64_t row0 = (int64_t)blockIdx.x * TROWS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // X tile (16-byte chunks), zero beyond nrows
    for (int c = tid; c < TROWS * INP / 8; c += 256) {
        const int r = (c * 8) / INP, k = (c * 8) % INP;
        bf16x8 v;
        if (row0 + r < nrows) {
            v = ld8(X + (row0 + r) * INP + k);
        } else {
            for (int q = 0; q < 8; ++q) v[q] = (bf16)0.0f;
        }
        *reinterpret_cast<bf16x8*>(xs + r * INP + k) = v;
    }
    __syncthreads();
    const int li = lane % 16, lg = lane / 16;
    const int rbase = 16 * wave;
    float ypart[4] = {0, 0, 0, 0};
    const bf16* in = xs;
    int K = INP;
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
        const bf16* Wl = W + (layer == 0 ? BOFF_W1 : layer == 1 ? BOFF_W2 : BOFF_W3);
        const float* bl = P + (layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : OFF_B3);
        bf16* out = hs[layer & 1];
        f32x4 acc[8];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) acc[nt] = (f32x4){0, 0, 0, 0};
        for (int ks = 0; ks < K / 32; ++ks) {
            const bf16x8 a = ld8(in + (rbase + li) * K + 32 * ks + 8 * lg);
#pragma unroll
            for (int nt = 0; nt < 8; ++nt) {
                const bf16x8 b = ld8(Wl + (16 * nt + li) * K + 32 * ks + 8 * lg);
                acc[nt] = mfma(a, b, acc[nt]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const int col = 16 * nt + li;
            const float bias = bl[col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rbase + 4 * lg + i;
                const float z = acc[nt][i] + bias;
                const float h = z * sigmoidf_(z);
                const bf16 hb = (bf16)h;
                out[r * HID + col] = hb;
                const int64_t grow = row0 + r;
                if (TRAIN && grow < nrows) {
                    const size_t o = (((size_t)layer * nmemb + m) * nrows + grow) * HID + col;
                    Z[o] = z;
                    Hs[o] = hb;
                }
                if (layer == 2) ypart[i] += (float)hb * P[OFF_W4 + col];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        in = out;
        K = HID;
    }
    // y = h3 . w4 + b4: reduce the 16 lanes of each row group
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = ypart[i];
        for (int s = 1; s < 16; s <<= 1) v += __shfl_xor(v, s, 64);
        const int64_t grow = row0 + rbase + 4 * lg + i;
        if (li == 0 && grow < nrows) y[(size_t)m * nrows + grow] = v + P[OFF_B4];
    }
}

// ---------------------------------------------------------------------------
// backward: grid (ntile, nmemb), block 256.  dy[m][row] is dL/dy.  Writes
// per-tile partial weight gradients (WGRAD) and/or input gradients dX[m][row][INP].

__device__ __forceinline__ float silu_d(float z) {
    const float s = sigmoidf_(z);
    return s * (1.0f + z * (1.0f - s));
}

template <bool WGRAD, bool XGRAD>
__global__ __launch_bounds__(256) void k_bwd(int64_t nrows, const bf16* __restrict__ X,
                                             const bf16* __restrict__ wb, const float* __restrict__ params,
                                             const float* __restrict__ dy, const float* __restrict__ Z,
                                             const bf16* __restrict__ Hs, float* __restrict__ part,
                                             float* __restrict__ dX) {
    __shared__ __attribute__((aligned(16))) bf16 g[TROWS * HID];     // dz, row-major [row][j]
    __shared__ __attribute__((aligned(16))) bf16 gT[HID * TROWS];    // dz transposed [j][row]
    __shared__ __attribute__((aligned(16))) bf16 hT[INP * TROWS];    // h_{l-1} transposed [k][row]
    __shared__ float dys[TROWS];
    const int m = blockIdx.y;
    const int nmemb = gridDim.y;
    const bf16* W = wb + (size_t)m * NBF;
    const float* P = params + (size_t)m * NPARAM;
    const int64_t row0 = (int64_t)blockIdx.x * TROWS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane % 16, lg = lane / 16;
    float* mypart = WGRAD ? part + ((size_t)m * gridDim.x + blockIdx.x) * NPARAM : nullptr;
    if (tid < TROWS) dys[tid] = (row0 + tid < nrows) ? dy[(size_t)m * nrows + row0 + tid] : 0.0f;
    __syncthreads();
    auto zval = [&](int layer, int r, int j) -> float {
        const int64_t grow = row0 + r;
        return grow < nrows ? Z[(((size_t)layer * nmemb + m) * nrows + grow) * HID + j] : 0.0f;
    };
    auto hval = [&](int layer, int r, int k) -> bf16 {
        const int64_t grow = row0 + r;
        if (grow >= nrows) return (bf16)0.0f;
        if (layer < 0) return X[grow * INP + k];
        return Hs[(((size_t)layer * nmemb + m) * nrows + grow) * HID + k];
    };
    // layer 4 and dz3 = dy * w4 * silu'(z3)
    for (int c = tid; c < TROWS * HID; c += 256) {
        const int r = c / HID, j = c % HID;
        const float dz = dys[r] * P[OFF_W4 + j] * silu_d(zval(2, r, j));
        const bf16 b = (bf16)dz;
        g[r * HID + j] = b;
        gT[j * TROWS + r] = b;
    }
    if (WGRAD) {
        // dw4[j] = sum_r dy h3[r][j]; db4 = sum_r dy
        for (int j = tid; j < HID; j += 256) {
            float s = 0.0f;
            for (int r = 0; r < TROWS; ++r) s += dys[r] * (float)hval(2, r, j);
            mypart[OFF_W4 + j] = s;
        }
        if (tid == 0) {
            float s = 0.0f;
            for (int r = 0; r < TROWS; ++r) s += dys[r];
            mypart[OFF_B4] = s;
        }
        for (int p = OFF_B4 + 1 + tid; p < NPARAM; p += 256) mypart[p] = 0.0f;  // padding
    }
    __syncthreads();
    for (int layer = 2; layer >= 0; --layer) {
        const int Kin = layer == 0 ? INP : HID;
        const int offW = layer == 0 ? OFF_W1 : layer == 1 ? OFF_W2 : OFF_W3;
        const int offB = layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : OFF_B3;
        const bf16* WT = W + (layer == 0 ? BOFF_T1 : layer == 1 ? BOFF_T2 : BOFF_T3);
        if (WGRAD) {
            // h_{l-1} transposed into LDS
            for (int c = tid; c < TROWS * Kin; c += 256) {
                const int r = c / Kin, k = c % Kin;
                hT[k * TROWS + r] = hval(layer - 1, r, k);
            }
            // db = sum over rows of dz
            for (int j = tid; j < HID; j += 256) {
                float s = 0.0f;
                for (int r = 0; r < TROWS; ++r) s += (float)g[r * HID + j];
                mypart[offB + j] = s;
            }
            __syncthreads();
            // dW[j][k] = sum_r dz[r][j] h[r][k]: tiles (jt, kt), 16x16, K = 64 rows
            const int ntk = Kin / 16;
            for (int t = wave; t < 8 * ntk; t += 4) {
                const int jt = t / ntk, kt = t % ntk;
                f32x4 acc = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < TROWS / 32; ++ks) {
                    const bf16x8 a = ld8(gT + (16 * jt + li) * TROWS + 32 * ks + 8 * lg);
                    const bf16x8 b = ld8(hT + (16 * kt + li) * TROWS + 32 * ks + 8 * lg);
                    acc = mfma(a, b, acc);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    mypart[offW + (16 * jt + 4 * lg + i) * Kin + 16 * kt + li] = acc[i];
            }
        }
        if (layer == 0 && !XGRAD) break;
        // dh_{l-1}[r][k] = sum_j dz[r][j] W[j][k]  (B operand from W^T [k][j])
        const int rbase = 16 * wave;
        const int ntk = Kin / 16;
        f32x4 acc[10];
#pragma unroll
        for (int kt = 0; kt < 10; ++kt) acc[kt] = (f32x4){0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < HID / 32; ++ks) {
            const bf16x8 a = ld8(g + (rbase + li) * HID + 32 * ks + 8 * lg);
#pragma unroll
            for (int kt = 0; kt < 10; ++kt) {
                if (kt < ntk) {
                    const bf16x8 b = ld8(WT + (16 * kt + li) * HID + 32 * ks + 8 * lg);
                    acc[kt] = mfma(a, b, acc[kt]);
                }
            }
        }
        __syncthreads();  // everyone done reading g before it is overwritten
        if (layer == 0) {
#pragma unroll
            for (int kt = 0; kt < 10; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t grow = row0 + rbase + 4 * lg + i;
                    if (kt < ntk && grow < nrows) dX[((size_t)m * nrows + grow) * INP + 16 * kt + li] = acc[kt][i];
                }
            break;
        }
#pragma unroll
        for (int kt = 0; kt < 8; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rbase + 4 * lg + i, k = 16 * kt + li;
                const float dz = acc[kt][i] * silu_d(zval(layer - 1, r, k));
                const bf16 b = (bf16)dz;
                g[r * HID + k] = b;
                gT[k * TROWS + r] = b;
            }
        __syncthreads();
    }
}

__global__ void k_fill(int64_t n, float v, float* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// sum the per-tile partials: grad[m][p] = sum_t part[m][t][p]
__global__ void k_reduce(int ntile, int nmemb, const float* __restrict__ part, float* __restrict__ grad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmemb * NPARAM) return;
    const int m = (int)(i / NPARAM), p = (int)(i % NPARAM);
    float s = 0.0f;
    for (int t = 0; t < ntile; ++t) s += part[((size_t)m * ntile + t) * NPARAM + p];
    grad[i] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master copy, then refresh the BF16 copies
__global__ void k_adam(int nmemb, float* __restrict__ params, const float* __restrict__ grad,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float wd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmemb * NPARAM) return;
    const float gr = grad[i] + wd * params[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gr;
    const float v = b2 * m2[i] + (1.0f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = v;
    params[i] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

__global__ void k_to_bf16(int nmemb, const float* __restrict__ params, bf16* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmemb * NBF) return;
    const int m = (int)(i / NBF), q = (int)(i % NBF);
    const float* P = params + (size_t)m * NPARAM;
    float v;
    if (q < BOFF_W2) {
        v = P[OFF_W1 + q];
    } else if (q < BOFF_W3) {
        v = P[OFF_W2 + (q - BOFF_W2)];
    } else if (q < BOFF_T1) {
        v = P[OFF_W3 + (q - BOFF_W3)];
    } else if (q < BOFF_T2) {
        const int t = q - BOFF_T1, k = t / HID, j = t % HID;  // T1[k][j] = W1[j][k]
        v = P[OFF_W1 + j * INP + k];
    } else if (q < BOFF_T3) {
        const int t = q - BOFF_T2, k = t / HID, j = t % HID;
        v = P[OFF_W2 + j * HID + k];
    } else {
        const int t = q - BOFF_T3, k = t / HID, j = t % HID;
        v = P[OFF_W3 + j * HID + k];
    }
    wb[i] = (bf16)v;
}

// MSE loss gradient on normalized targets: dy = 2 (y - t) / n; the loss of
// member blockIdx.y as wave sums, one atomic add per wave (the loss is a
// diagnostic: its summation order is not fixed)
__global__ void k_mse(int64_t nrows, int nmemb, const float* __restrict__ y, const float* __restrict__ t,
                      float* __restrict__ dy, float* __restrict__ loss) {
    const int m = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (r < nrows) {
        const int64_t i = (int64_t)m * nrows + r;
        const float e = y[i] - t[r];
        dy[i] = 2.0f * e / (float)nrows;
        s = e * e / (float)nrows;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(&loss[m], s);
}

// ---------------------------------------------------------------------------
// host

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t c) {
        if (c <= n) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        HIP_TRY(hipMalloc(&p, std::max<size_t>(c, 1) * sizeof(T)));
        n = c;
        return 0;
    }
};

struct acemlp_ctx {
    int device = 0, nmemb = 0, nin = 0;
    int64_t step = 0;
    Buf<float> params, grad, m1, m2, part, Z, dy, loss, y;
    Buf<bf16> wb, Hs;
};

extern "C" int acemlp_nparam(void) { return NPARAM; }
extern "C" int acemlp_input_width(void) { return INP; }
extern "C" const char* acemlp_last_error(void) { return g_err.c_str(); }

extern "C" int acemlp_create(int nmemb, int nin, int device, uint64_t seed, acemlp_ctx** out) {
    if (!out) return set_err("null out");
    *out = nullptr;
    if (nmemb < 1 || nmemb > 16) return set_err("nmemb must be in [1, 16]");
    if (nin < 1 || nin > INP) return set_err("input width must be in [1, 160]");
    HIP_TRY(hipSetDevice(device));
    acemlp_ctx* c = new acemlp_ctx();
    c->device = device;
    c->nmemb = nmemb;
    c->nin = nin;
    // He-uniform initialization; padded input columns are zero
    std::vector<float> h((size_t)nmemb * NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    for (int m = 0; m < nmemb; ++m) {
        float* P = h.data() + (size_t)m * NPARAM;
        auto fill = [&](int off, int rows, int cols, int fan_in, int valid_cols) {
            std::uniform_real_distribution<float> U(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
            for (int r = 0; r < rows; ++r)
                for (int k = 0; k < cols; ++k) P[off + r * cols + k] = k < valid_cols ? U(rng) : 0.0f;
        };
        fill(OFF_W1, HID, INP, nin, nin);
        fill(OFF_W2, HID, HID, HID, HID);
        fill(OFF_W3, HID, HID, HID, HID);
        fill(OFF_W4, 1, HID, HID, HID);
    }
    if (c->params.ensure(h.size()) || c->grad.ensure(h.size()) || c->m1.ensure(h.size()) ||
        c->m2.ensure(h.size()) || c->wb.ensure((size_t)nmemb * NBF) || c->loss.ensure(nmemb)) {
        delete c;
        return -1;
    }
    HIP_TRY(hipMemcpy(c->params.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(c->m1.p, 0, h.size() * sizeof(float)));
    HIP_TRY(hipMemset(c->m2.p, 0, h.size() * sizeof(float)));
    const int64_t nb = (int64_t)nmemb * NBF;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, 0, nmemb, c->params.p, c->wb.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    *out = c;
    return 0;
}

extern "C" void acemlp_destroy(acemlp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    delete c;
}

extern "C" int acemlp_get_params(acemlp_ctx* c, float* host, int64_t n) {
    if (!c || !host || n != (int64_t)c->nmemb * NPARAM) return set_err("acemlp_get_params: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(host, c->params.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int acemlp_set_params(acemlp_ctx* c, const float* host, int64_t n) {
    if (!c || !host || n != (int64_t)c->nmemb * NPARAM) return set_err("acemlp_set_params: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(c->params.p, host, n * sizeof(float), hipMemcpyHostToDevice));
    const int64_t nb = (int64_t)c->nmemb * NBF;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, 0, c->nmemb, c->params.p, c->wb.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

static int ntile_of(int64_t n) { return (int)((n + TROWS - 1) / TROWS); }

// y[m][row] for bf16 X[row][160] (normalized features, zero padded)
extern "C" int acemlp_predict(acemlp_ctx* c, int64_t nrows, const void* X, float* y, void* stream) {
    if (!c || nrows < 0 || (nrows > 0 && (!X || !y))) return set_err("acemlp_predict: bad arguments");
    if (nrows == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_fwd<false>, dim3(ntile_of(nrows), c->nmemb), dim3(256), 0, (hipStream_t)stream, nrows,
                       (const bf16*)X, c->wb.p, c->params.p, y, (float*)nullptr, (bf16*)nullptr);
    HIP_TRY(hipGetLastError());
    return 0;
}

// forward with stored activations + backward.  dy given (device [m][row]).
static int fwd_bwd(acemlp_ctx* c, int64_t nrows, const bf16* X, const float* dy_ext, const float* target,
                   bool wgrad, float* dX, hipStream_t st) {
    const int nt = ntile_of(nrows);
    if (c->Z.ensure((size_t)3 * c->nmemb * nrows * HID) || c->Hs.ensure((size_t)3 * c->nmemb * nrows * HID) ||
        c->y.ensure((size_t)c->nmemb * nrows) || c->dy.ensure((size_t)c->nmemb * nrows))
        return -1;
    if (wgrad && c->part.ensure((size_t)c->nmemb * nt * NPARAM)) return -1;
    hipLaunchKernelGGL(k_fwd<true>, dim3(nt, c->nmemb), dim3(256), 0, st, nrows, X, c->wb.p, c->params.p, c->y.p,
                       c->Z.p, c->Hs.p);
    HIP_TRY(hipGetLastError());
    const float* dy = dy_ext;
    if (target) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (int64_t)c->nmemb, 0.0f, c->loss.p);
        hipLaunchKernelGGL(k_mse, dim3((unsigned)((nrows + 255) / 256), (unsigned)c->nmemb), dim3(256), 0, st, nrows,
                           c->nmemb, c->y.p, target, c->dy.p, c->loss.p);
        HIP_TRY(hipGetLastError());
        dy = c->dy.p;
    }
    if (wgrad && dX)
        hipLaunchKernelGGL((k_bwd<true, true>), dim3(nt, c->nmemb), dim3(256), 0, st, nrows, X, c->wb.p,
                           c->params.p, dy, c->Z.p, c->Hs.p, c->part.p, dX);
    else if (wgrad)
        hipLaunchKernelGGL((k_bwd<true, false>), dim3(nt, c->nmemb), dim3(256), 0, st, nrows, X, c->wb.p,
                           c->params.p, dy, c->Z.p, c->Hs.p, c->part.p, (float*)nullptr);
    else
        hipLaunchKernelGGL((k_bwd<false, true>), dim3(nt, c->nmemb), dim3(256), 0, st, nrows, X, c->wb.p,
                           c->params.p, dy, c->Z.p, c->Hs.p, (float*)nullptr, dX);
    HIP_TRY(hipGetLastError());
    if (wgrad) {
        const int64_t n = (int64_t)c->nmemb * NPARAM;
        hipLaunchKernelGGL(k_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nt, c->nmemb, c->part.p,
                           c->grad.p);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// gradient of the MSE loss (normalized targets t[row]) w.r.t. the parameters
extern "C" int acemlp_loss_grad(acemlp_ctx* c, int64_t nrows, const void* X, const float* t, float* grad_host,
                                float* loss_host) {
    if (!c || nrows < 1 || !X || !t) return set_err("acemlp_loss_grad: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    if (fwd_bwd(c, nrows, (const bf16*)X, nullptr, t, true, nullptr, 0)) return -1;
    HIP_TRY(hipDeviceSynchronize());
    if (grad_host)
        HIP_TRY(hipMemcpy(grad_host, c->grad.p, sizeof(float) * c->nmemb * NPARAM, hipMemcpyDeviceToHost));
    if (loss_host) HIP_TRY(hipMemcpy(loss_host, c->loss.p, sizeof(float) * c->nmemb, hipMemcpyDeviceToHost));
    return 0;
}

// one Adam step on the MSE loss of the batch
extern "C" int acemlp_train_step(acemlp_ctx* c, int64_t nrows, const void* X, const float* t, float lr,
                                 float weight_decay, float* loss_host, void* stream) {
    if (!c || nrows < 1 || !X || !t) return set_err("acemlp_train_step: bad arguments");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    if (fwd_bwd(c, nrows, (const bf16*)X, nullptr, t, true, nullptr, st)) return -1;
    c->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, (float)c->step), bc2 = 1.0f - std::pow(b2, (float)c->step);
    const int64_t n = (int64_t)c->nmemb * NPARAM;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c->nmemb, c->params.p,
                       c->grad.p, c->m1.p, c->m2.p, lr, b1, b2, 1e-8f, bc1, bc2, weight_decay);
    const int64_t nb = (int64_t)c->nmemb * NBF;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, c->nmemb, c->params.p,
                       c->wb.p);
    HIP_TRY(hipGetLastError());
    if (loss_host) {
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipMemcpy(loss_host, c->loss.p, sizeof(float) * c->nmemb, hipMemcpyDeviceToHost));
    }
    return 0;
}

// y[m][row] and dX[m][row][160] = d y_m / d X (for forces through the surrogate)
extern "C" int acemlp_predict_grad(acemlp_ctx* c, int64_t nrows, const void* X, float* y, float* dX, void* stream) {
    if (!c || nrows < 0 || (nrows > 0 && (!X || !y || !dX))) return set_err("acemlp_predict_grad: bad arguments");
    if (nrows == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    if (c->dy.ensure((size_t)c->nmemb * nrows)) return -1;
    // dy = 1 for every member (a fill kernel, no memset)
    const int64_t n = (int64_t)c->nmemb * nrows;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, 1.0f, c->dy.p);
    HIP_TRY(hipGetLastError());
    if (fwd_bwd(c, nrows, (const bf16*)X, c->dy.p, nullptr, false, dX, st)) return -1;
    HIP_TRY(hipMemcpyAsync(y, c->y.p, sizeof(float) * c->nmemb * nrows, hipMemcpyDeviceToDevice, st));
    return 0;
}

extern "C" int acemlp_layout_probe(float* host16x16) {
    float* d = nullptr;
    HIP_TRY(hipMalloc(&d, 256 * sizeof(float)));
    hipLaunchKernelGGL(k_layout_probe, dim3(1), dim3(64), 0, 0, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(host16x16, d, 256 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(hipGetErrorString(e));
    return 0;
}

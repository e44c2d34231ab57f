// arbo_surrogate.hip — the surrogate MLP on gfx950: matrix products on the
// bfloat16 MFMA units with float32 accumulation, fused bias / GELU / GELU
// gradient / bias-gradient epilogues, the Poisson loss, and Adam.
//
// One kernel does every matrix product of the forward and backward passes:
// C[i, j] = sum_k A(i, k) B(j, k), each operand float32 in memory with its
// own strides (one of them 1), so X W^T (forward), dZ W (input gradients)
// and dZ^T X (weight gradients over the batch) are the same code.
//  * A workgroup of four wave64s computes a 64 x 64 block of C; each wave one
//    32 x 32 tile with v_mfma_f32_32x32x16_bf16 (16-deep k-steps, float32
//    accumulators, 16 per lane).
//  * The block's A and B panels, 64 rows x 64 k each, are staged through
//    LDS: a thread loads 16 consecutive elements along the operand's
//    contiguous dimension (16-byte loads where aligned), splits each float
//    into two bfloat16 halves, hi = bf16(x) and lo = bf16(x - hi), and
//    stores both; a lane then reads its 8 k-values as one 16-byte LDS read
//    per half.  The next panels' loads are issued before the current
//    panels' MFMAs, so their latency overlaps the arithmetic.
//  * Each k-step issues three MFMAs, hi*hi + hi*lo + lo*hi: the products
//    carry about 16 significant bits of every operand (the bf16x3 split),
//    so gradients hold to float32-like accuracy while every multiply runs
//    on the bfloat16 matrix cores.
//  * The weight-gradient product appends a column of ones to its B operand,
//    so the bias gradient (the batch sum of dZ) comes out of the same launch,
//    and the weight gradients of all layers, independent once the input
//    gradients are known, share one launch (k_gemm_group): a training step
//    is then 5 forward products, 4 input-gradient products and one grouped
//    launch, where each launch costs about 5 us whatever its size.
// Parameters stay float32 and are read by the products directly, so an
// Adam step needs no repacking.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "arbo_c.h"
#include "arbo_rng.h"
#include "arbo_surrogate.h"

namespace arbo {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kBM = 64, kBN = 64, kBK = 64;   // block tile; 4 waves, one 32 x 32 MFMA tile each
constexpr uint32_t kLdsK = kBK + 8;                 // padded LDS row of a panel (bf16): 144 bytes
constexpr uint32_t kPer = kBM * kBK / 256;          // panel elements per thread: 16

// float32 -> bfloat16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum : uint32_t {
    EPI_STORE = 0,       // outf = acc (with ones: column N - 1 to outb)
    EPI_BIAS = 1,        // outf = acc + bias
    EPI_BIAS_GELU = 2,   // outf = z = acc + bias; outh = gelu(z)
    EPI_GELU_GRAD = 3,   // outf = acc * gelu'(zprev)
    EPI_GELU_ONLY = 4    // outh = gelu(acc + bias) (forward passes that keep no z)
};

struct GemmArgs {
    uint32_t M, N, K, epi;
    const float* A;
    uint64_t sAi, sAk;      // A(i, k) = A[i sAi + k sAk]; sAk == 1 (k-contiguous) or sAi == 1
    const float* B;
    uint64_t sBj, sBk;      // B(j, k) = B[j sBj + k sBk]; sBk == 1 (k-contiguous) or sBj == 1
    uint32_t vecA, vecB;    // 1: 16-byte loads along the contiguous dimension are aligned
    uint32_t ones;          // 1: B(N - 1, k) = 1 for every k (its column of C is a row sum of A)
    const float* bias;
    const float* zprev;
    uint32_t ldz;
    float* outf;
    uint32_t ldf;
    float* outh;
    uint32_t ldh;
    float* outb;            // with ones: column N - 1 of C goes to outb[i]
};

struct Panel {
    uint16_t hi[kBM][kLdsK];
    uint16_t lo[kBM][kLdsK];
};

// One thread's share of a 64 x 64 panel (operand rows [r0, r0 + 64), k in
// [k0, k0 + 64)): 16 consecutive elements along the operand's contiguous
// dimension.  rows: the operand's rows that exist in memory; with `ones`
// row `rows` is all ones.
__device__ inline void load_panel(const float* p, uint64_t s_row, uint64_t s_k, uint32_t rows, uint32_t K,
                                  uint32_t r0, uint32_t k0, bool vec, bool ones, uint32_t tid, float* v) {
    const uint32_t a = tid >> 2, c = (tid & 3u) * kPer;   // k-contiguous: row a, k c..; else k a, rows c..
    if (s_k == 1) {
        const uint32_t gr = r0 + a, gk = k0 + c;
        if (gr < rows) {
            const float* q = p + (uint64_t)gr * s_row + gk;
            if (vec && gk + kPer <= K) {
#pragma unroll
                for (uint32_t e = 0; e < kPer; e += 4) {
                    const float4 f = reinterpret_cast<const float4*>(q + e)[0];
                    v[e] = f.x; v[e + 1] = f.y; v[e + 2] = f.z; v[e + 3] = f.w;
                }
            } else {
#pragma unroll
                for (uint32_t e = 0; e < kPer; ++e) v[e] = gk + e < K ? q[e] : 0.0f;
            }
        } else {
#pragma unroll
            for (uint32_t e = 0; e < kPer; ++e) v[e] = (ones && gr == rows && gk + e < K) ? 1.0f : 0.0f;
        }
    } else {   // s_row == 1
        const uint32_t gk = k0 + a, gr = r0 + c;
        if (gk < K) {
            const float* q = p + (uint64_t)gk * s_k + gr;
            if (vec && gr + kPer <= rows) {
#pragma unroll
                for (uint32_t e = 0; e < kPer; e += 4) {
                    const float4 f = reinterpret_cast<const float4*>(q + e)[0];
                    v[e] = f.x; v[e + 1] = f.y; v[e + 2] = f.z; v[e + 3] = f.w;
                }
            } else {
#pragma unroll
                for (uint32_t e = 0; e < kPer; ++e)
                    v[e] = gr + e < rows ? q[e] : ((ones && gr + e == rows) ? 1.0f : 0.0f);
            }
        } else {
#pragma unroll
            for (uint32_t e = 0; e < kPer; ++e) v[e] = 0.0f;
        }
    }
}

__device__ inline uint32_t pack2(uint16_t a, uint16_t b) { return (uint32_t)a | ((uint32_t)b << 16); }

// Splits a thread's share into bfloat16 halves, hi = bf16(x) and lo =
// bf16(x - hi), and stores both into the panel in [row][k] order.
__device__ inline void store_panel(const float* v, bool kcontig, uint32_t tid, Panel& P) {
    const uint32_t a = tid >> 2, c = (tid & 3u) * kPer;
    uint16_t h[kPer], l[kPer];
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) {
        h[e] = f2bf(v[e]);
        l[e] = f2bf(v[e] - bf2f(h[e]));
    }
    if (kcontig) {
#pragma unroll
        for (uint32_t e = 0; e < kPer; e += 8) {
            *reinterpret_cast<uint4*>(&P.hi[a][c + e]) =
                make_uint4(pack2(h[e], h[e + 1]), pack2(h[e + 2], h[e + 3]), pack2(h[e + 4], h[e + 5]),
                           pack2(h[e + 6], h[e + 7]));
            *reinterpret_cast<uint4*>(&P.lo[a][c + e]) =
                make_uint4(pack2(l[e], l[e + 1]), pack2(l[e + 2], l[e + 3]), pack2(l[e + 4], l[e + 5]),
                           pack2(l[e + 6], l[e + 7]));
        }
    } else {
#pragma unroll
        for (uint32_t e = 0; e < kPer; ++e) {
            P.hi[c + e][a] = h[e];
            P.lo[c + e][a] = l[e];
        }
    }
}

// MFMA operand layout of v_mfma_f32_32x32x16_bf16: lane l holds A(l % 32,
// 8 (l / 32) + v) and B(8 (l / 32) + v, l % 32) for v = 0..7; accumulator r
// of lane l is D(8 (r / 4) + 4 (l / 32) + r % 4, l % 32).
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, uint32_t bx, uint32_t by, Panel& sa, Panel& sb) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t i0 = by * kBM, j0 = bx * kBN;
    const uint32_t wi = (wave >> 1) * 32u, wj = (wave & 1u) * 32u;
    const uint32_t rowsB = g.ones ? g.N - 1 : g.N;
    const bool ka = g.sAk == 1, kb = g.sBk == 1;
    f32x16 acc = {};
    // the next panels' elements are loaded while the current ones multiply
    float ra[kPer], rb[kPer];
    load_panel(g.A, g.sAi, g.sAk, g.M, g.K, i0, 0, g.vecA != 0, false, tid, ra);
    load_panel(g.B, g.sBj, g.sBk, rowsB, g.K, j0, 0, g.vecB != 0, g.ones != 0, tid, rb);
    for (uint32_t k0 = 0; k0 < g.K; k0 += kBK) {
        store_panel(ra, ka, tid, sa);
        store_panel(rb, kb, tid, sb);
        __syncthreads();
        if (k0 + kBK < g.K) {
            load_panel(g.A, g.sAi, g.sAk, g.M, g.K, i0, k0 + kBK, g.vecA != 0, false, tid, ra);
            load_panel(g.B, g.sBj, g.sBk, rowsB, g.K, j0, k0 + kBK, g.vecB != 0, g.ones != 0, tid, rb);
        }
#pragma unroll
        for (uint32_t kk = 0; kk < kBK; kk += 16) {
            const uint32_t kq = kk + 8u * (lane >> 5);
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&sa.hi[wi + (lane & 31u)][kq]);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(&sa.lo[wi + (lane & 31u)][kq]);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&sb.hi[wj + (lane & 31u)][kq]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&sb.lo[wj + (lane & 31u)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const uint32_t j = j0 + wj + (lane & 31u);
    if (j >= g.N) return;
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) {
        const uint32_t i = i0 + wi + 8u * (r >> 2) + 4u * (lane >> 5) + (r & 3u);
        if (i >= g.M) continue;
        const float v = acc[r];
        switch (g.epi) {
            case EPI_STORE:
                if (g.ones && j == g.N - 1) g.outb[i] = v;
                else g.outf[(size_t)i * g.ldf + j] = v;
                break;
            case EPI_BIAS:
                g.outf[(size_t)i * g.ldf + j] = v + g.bias[j];
                break;
            case EPI_BIAS_GELU: {
                const float z = v + g.bias[j];
                g.outf[(size_t)i * g.ldf + j] = z;
                g.outh[(size_t)i * g.ldh + j] = gelu(z);
                break;
            }
            case EPI_GELU_GRAD:
                g.outf[(size_t)i * g.ldf + j] = v * gelu_grad(g.zprev[(size_t)i * g.ldz + j]);
                break;
            default:
                g.outh[(size_t)i * g.ldh + j] = gelu(v + g.bias[j]);
                break;
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ Panel sa, sb;
    gemm_tile(g, blockIdx.x, blockIdx.y, sa, sb);
}

// Independent products in one launch: product z = blockIdx.z of the group,
// each over the tiles it has (a workgroup past them leaves at once).  The
// weight gradients of every layer are independent once the input gradients
// are known, so a backward pass runs them together.
struct GemmGroup {
    GemmArgs g[ARBO_SUR_MAX_LAYERS];
};

__global__ __launch_bounds__(256) void k_gemm_group(GemmGroup gg) {
    __shared__ Panel sa, sb;
    const GemmArgs& g = gg.g[blockIdx.z];
    if (blockIdx.x * kBN >= g.N || blockIdx.y * kBM >= g.M) return;   // workgroup-uniform
    gemm_tile(g, blockIdx.x, blockIdx.y, sa, sb);
}

// rows idx[r] of x [*, w] to out [r, w]
__global__ void k_gather(const float* __restrict__ x, const uint32_t* __restrict__ idx, float* __restrict__ out,
                         uint32_t rows, uint32_t w) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)rows * w) return;
    const uint32_t r = (uint32_t)(t / w), c = (uint32_t)(t % w);
    out[t] = x[(uint64_t)idx[r] * w + c];
}

// Poisson loss terms of the output layer: dF = (exp(f) - y) * scale; the
// loss's partial sums per workgroup into part[blockIdx.x] and, for a running
// epoch sum, into *acc
// With `rows` the targets are rows rows[r] of y (width w): the minibatch's
// rows gathered on the fly.
__global__ __launch_bounds__(256) void k_poisson(const float* __restrict__ f, const float* __restrict__ y,
                                                 const uint32_t* __restrict__ rows, uint32_t w, uint64_t n,
                                                 float scale, float* __restrict__ df, double* __restrict__ part,
                                                 double* __restrict__ acc) {
    __shared__ double s[256];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float yi = rows ? y[(uint64_t)rows[i / w] * w + i % w] : y[i];
        const float e = expf(f[i]);
        df[i] = (e - yi) * scale;
        l = (double)e - (double)yi * (double)f[i];
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s[0];
        if (acc) atomicAdd(acc, s[0]);
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint64_t n, float lr, float c1, float c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ void k_zero_f32(float* __restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

template <class T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
};

// Activations of arbo_sur_forward_noalloc: two ping-pong buffers of
// kScratchElems float32 values (8 MiB in all) that every GPU the library
// runs on holds until the process ends.  Host code never uses the symbol
// itself (on the host it only names the variable): it takes each device's
// address once, from hipGetSymbolAddress with that device current, when a
// surrogate is created there.
constexpr size_t kScratchElems = size_t(1) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[2 * kScratchElems];

std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on `device`, which must be current.
hipError_t scratch_on(int device, float** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess) return st;
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace
}  // namespace arbo

using namespace arbo;

struct ArboSurrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<uint32_t> w;        // widths
    uint32_t L = 0, max_batch = 0;
    std::vector<size_t> woff, boff; // offsets of W_l and b_l in the flat parameter array
    size_t n_params = 0;
    Buf<float> P, G, Mo, Vo;        // parameters, gradients, Adam moments
    std::vector<Buf<float>> Z;      // pre-activations [B, w_{l+1}]
    std::vector<Buf<float>> H;      // layer inputs [B, w_l] (H[0]: the rows' inputs)
    std::vector<Buf<float>> dZ;     // gradients of the pre-activations [B, w_{l+1}]
    Buf<float> yin, dxin;
    Buf<double> part;               // the loss's per-workgroup sums (max_batch x outputs / 256)
    Buf<double> epoch_sum;          // the loss summed over one training epoch
    Buf<float> data_x, data_y;
    Buf<uint32_t> idx;
    float* scratch = nullptr;       // this device's no-allocation activation buffers
    uint64_t adam_t = 0;
    std::string err;
};

namespace {

int fail(ArboSurrogate* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

#define SUR_TRY(s, call)                                                                    \
    do {                                                                                    \
        hipError_t _st = (call);                                                            \
        if (_st != hipSuccess)                                                              \
            return fail((s), ARBO_E_DEVICE, std::string(#call ": ") + hipGetErrorString(_st)); \
    } while (0)

template <class T>
int alloc(ArboSurrogate* s, Buf<T>& b, size_t n) {
    if (b.p && b.n >= n) return ARBO_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    if (n == 0) return ARBO_OK;
    if (hipMalloc((void**)&b.p, n * sizeof(T)) != hipSuccess) {
        b.p = nullptr;
        return fail(s, ARBO_E_ALLOC, "hipMalloc failed");
    }
    b.n = n;
    return ARBO_OK;
}

template <class T>
void release(Buf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

uint32_t blocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

void set_vec(GemmArgs& g) {
    g.vecA = aligned16(g.A) && ((g.sAk == 1 ? g.sAi : g.sAk) % 4 == 0);
    g.vecB = aligned16(g.B) && ((g.sBk == 1 ? g.sBj : g.sBk) % 4 == 0);
}

// C[M, N] = sum_k A(i, k) B(j, k) with the epilogue of g.epi.
hipError_t gemm(hipStream_t st, GemmArgs g) {
    set_vec(g);
    dim3 grid(blocks(g.N, kBN), blocks(g.M, kBM));
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

// n independent products in one launch.
hipError_t gemm_group(hipStream_t st, GemmGroup gg, uint32_t n) {
    uint32_t bx = 1, by = 1;
    for (uint32_t z = 0; z < n; ++z) {
        set_vec(gg.g[z]);
        bx = std::max(bx, blocks(gg.g[z].N, kBN));
        by = std::max(by, blocks(gg.g[z].M, kBM));
    }
    hipLaunchKernelGGL(k_gemm_group, dim3(bx, by, n), dim3(256), 0, st, gg);
    return hipGetLastError();
}

GemmArgs args(uint32_t M, uint32_t N, uint32_t K, uint32_t epi) {
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.epi = epi;
    return g;
}

// Forward pass of B rows in H[0]: z of every layer in Z, the hidden layers'
// activations in H; the output layer's z in Z[L-1].
int forward(ArboSurrogate* s, uint32_t B) {
    for (uint32_t l = 0; l < s->L; ++l) {
        const uint32_t K = s->w[l], N = s->w[l + 1];
        const bool last = l + 1 == s->L;
        GemmArgs g = args(B, N, K, last ? EPI_BIAS : EPI_BIAS_GELU);
        g.A = s->H[l].p;
        g.sAi = K;
        g.sAk = 1;
        g.B = s->P.p + s->woff[l];   // W_l [N, K]
        g.sBj = K;
        g.sBk = 1;
        g.bias = s->P.p + s->boff[l];
        g.outf = s->Z[l].p;
        g.ldf = N;
        if (!last) {
            g.outh = s->H[l + 1].p;
            g.ldh = N;
        }
        SUR_TRY(s, gemm(s->stream, g));
    }
    return ARBO_OK;
}

// Backward pass from dZ[L-1]: the input-gradient chain dZ[l-1] = (dZ[l] W_l)
// * gelu'(Z[l-1]) down to layer 1 (and to the inputs, into dxin, when
// want_x); then, when want_w, the weight and bias gradients of every layer
// in one launch ([dW | db] = dZ^T [X | 1]: the bias gradient is the ones
// column), since they are independent once the dZ are known.
int backward(ArboSurrogate* s, uint32_t B, bool want_w, bool want_x) {
    for (int l = (int)s->L - 1; l >= (want_x ? 0 : 1); --l) {
        const uint32_t K = s->w[l], N = s->w[l + 1];
        GemmArgs g = args(B, K, N, l > 0 ? EPI_GELU_GRAD : EPI_STORE);   // dX = dZ W
        g.A = s->dZ[l].p;
        g.sAi = N;
        g.sAk = 1;
        g.B = s->P.p + s->woff[l];   // B(j, k) = W[k, j]
        g.sBj = 1;
        g.sBk = K;
        if (l > 0) {
            g.zprev = s->Z[l - 1].p;
            g.ldz = K;
            g.outf = s->dZ[l - 1].p;
            g.ldf = K;
        } else {
            g.outf = s->dxin.p;
            g.ldf = K;
        }
        SUR_TRY(s, gemm(s->stream, g));
    }
    if (want_w) {
        GemmGroup gg{};
        for (uint32_t l = 0; l < s->L; ++l) {
            const uint32_t K = s->w[l], N = s->w[l + 1];
            GemmArgs& g = gg.g[l];
            g = args(N, K + 1, B, EPI_STORE);
            g.A = s->dZ[l].p;
            g.sAi = 1;
            g.sAk = N;
            g.B = s->H[l].p;
            g.sBj = 1;
            g.sBk = K;
            g.ones = 1;
            g.outf = s->G.p + s->woff[l];
            g.ldf = K;
            g.outb = s->G.p + s->boff[l];
        }
        SUR_TRY(s, gemm_group(s->stream, gg, s->L));
    }
    return ARBO_OK;
}

int upload_x(ArboSurrogate* s, uint32_t n, const float* x) {
    SUR_TRY(s, hipMemcpyAsync(s->H[0].p, x, (size_t)n * s->w[0] * 4, hipMemcpyHostToDevice, s->stream));
    return ARBO_OK;
}

// Poisson loss and output gradient for B rows with targets in `y` (device);
// the loss sum is returned through *loss (synchronises).
int poisson(ArboSurrogate* s, uint32_t B, const float* y, float scale, double* loss, double* acc = nullptr,
            const uint32_t* rows = nullptr) {
    const uint32_t out = s->w[s->L];
    const uint64_t n = (uint64_t)B * out;
    const uint32_t nb = blocks(n, 256);
    hipLaunchKernelGGL(k_poisson, dim3(nb), dim3(256), 0, s->stream, s->Z[s->L - 1].p, y, rows, out, n, scale,
                       s->dZ[s->L - 1].p, s->part.p, acc);
    SUR_TRY(s, hipGetLastError());
    if (loss) {
        std::vector<double> h(nb);
        SUR_TRY(s, hipMemcpyAsync(h.data(), s->part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        SUR_TRY(s, hipStreamSynchronize(s->stream));
        double t = 0.0;
        for (double v : h) t += v;
        *loss = t;
    }
    return ARBO_OK;
}

int adam_step(ArboSurrogate* s, float lr) {
    s->adam_t++;
    const float c1 = 1.0f - std::pow(0.9f, (float)s->adam_t), c2 = 1.0f - std::pow(0.999f, (float)s->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(s->n_params, 256)), dim3(256), 0, s->stream, s->P.p, s->G.p, s->Mo.p,
                       s->Vo.p, (uint64_t)s->n_params, lr, c1, c2);
    SUR_TRY(s, hipGetLastError());
    return ARBO_OK;
}

}  // namespace

extern "C" {

ARBO_API int arbo_sur_create(int32_t device, uint32_t n_widths, const uint32_t* widths, uint32_t max_batch,
                             uint64_t seed, ArboSurrogate** out) {
    if (!out) return ARBO_E_INVALID;
    *out = nullptr;
    if (!widths || n_widths < 2 || n_widths > ARBO_SUR_MAX_LAYERS + 1 || max_batch == 0 || max_batch > (1u << 20))
        return ARBO_E_INVALID;
    for (uint32_t i = 0; i < n_widths; ++i)
        if (widths[i] == 0 || widths[i] > ARBO_SUR_MAX_WIDTH) return ARBO_E_INVALID;
    ArboSurrogate* s = new (std::nothrow) ArboSurrogate();
    if (!s) return ARBO_E_ALLOC;
    s->device = device;
    s->w.assign(widths, widths + n_widths);
    s->L = n_widths - 1;
    s->max_batch = max_batch;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        scratch_on(device, &s->scratch) != hipSuccess) {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return ARBO_E_DEVICE;
    }
    size_t off = 0;
    for (uint32_t l = 0; l < s->L; ++l) {
        s->woff.push_back(off);
        off += (size_t)s->w[l + 1] * s->w[l];
        s->boff.push_back(off);
        off += s->w[l + 1];
    }
    s->n_params = off;
    s->Z.resize(s->L);
    s->H.resize(s->L);
    s->dZ.resize(s->L);
    int rc = alloc(s, s->P, off);
    if (!rc) rc = alloc(s, s->G, off);
    if (!rc) rc = alloc(s, s->Mo, off);
    if (!rc) rc = alloc(s, s->Vo, off);
    for (uint32_t l = 0; l < s->L && !rc; ++l) {
        rc = alloc(s, s->Z[l], (size_t)max_batch * s->w[l + 1]);
        if (!rc) rc = alloc(s, s->H[l], (size_t)max_batch * s->w[l]);
        if (!rc) rc = alloc(s, s->dZ[l], (size_t)max_batch * s->w[l + 1]);
    }
    if (!rc) rc = alloc(s, s->dxin, (size_t)max_batch * s->w[0]);
    if (!rc) rc = alloc(s, s->yin, (size_t)max_batch * s->w[s->L]);
    if (!rc) rc = alloc(s, s->part, blocks((uint64_t)max_batch * s->w[s->L], 256));
    if (!rc) rc = alloc(s, s->epoch_sum, 1);
    if (rc) {
        arbo_sur_destroy(s);
        return rc;
    }
    // He-normal weights from Philox (Box-Muller), zero biases
    std::vector<float> h(off, 0.0f);
    const Key key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t l = 0; l < s->L; ++l) {
        const double sd = std::sqrt(2.0 / s->w[l]);
        const size_t n = (size_t)s->w[l + 1] * s->w[l];
        for (size_t i = 0; i < n; i += 2) {
            const U4 r = draw(key, l, (uint32_t)(i >> 1), (uint32_t)((uint64_t)i >> 33), 0x5u);
            const double u1 = 1.0 - u53(r.x, r.y), u2 = u53(r.z, r.w);
            const double rad = std::sqrt(-2.0 * std::log(u1));
            h[s->woff[l] + i] = (float)(sd * rad * std::cos(6.283185307179586 * u2));
            if (i + 1 < n) h[s->woff[l] + i + 1] = (float)(sd * rad * std::sin(6.283185307179586 * u2));
        }
    }
    rc = arbo_sur_set_params(s, h.data());
    if (rc) {
        arbo_sur_destroy(s);
        return rc;
    }
    *out = s;
    return ARBO_OK;
}

ARBO_API int arbo_sur_n_params(const ArboSurrogate* s, uint64_t* out) {
    if (!s || !out) return ARBO_E_INVALID;
    *out = s->n_params;
    return ARBO_OK;
}

ARBO_API int arbo_sur_shape(const ArboSurrogate* s, uint32_t* inputs, uint32_t* outputs) {
    if (!s || !inputs || !outputs) return ARBO_E_INVALID;
    *inputs = s->w.front();
    *outputs = s->w.back();
    return ARBO_OK;
}

ARBO_API int arbo_sur_get_params(ArboSurrogate* s, float* out) {
    if (!s || !out) return ARBO_E_INVALID;
    SUR_TRY(s, hipSetDevice(s->device));
    SUR_TRY(s, hipMemcpyAsync(out, s->P.p, s->n_params * 4, hipMemcpyDeviceToHost, s->stream));
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API int arbo_sur_set_params(ArboSurrogate* s, const float* in) {
    if (!s || !in) return ARBO_E_INVALID;
    SUR_TRY(s, hipSetDevice(s->device));
    SUR_TRY(s, hipMemcpyAsync(s->P.p, in, s->n_params * 4, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks(s->n_params, 256)), dim3(256), 0, s->stream, s->Mo.p,
                       (uint64_t)s->n_params);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks(s->n_params, 256)), dim3(256), 0, s->stream, s->Vo.p,
                       (uint64_t)s->n_params);
    SUR_TRY(s, hipGetLastError());
    s->adam_t = 0;
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API int arbo_sur_forward(ArboSurrogate* s, uint32_t n, const float* x, float* out) {
    if (!s || !x || !out || n == 0) return ARBO_E_INVALID;
    if (n > s->max_batch) return fail(s, ARBO_E_CAPACITY, "more rows than max_batch");
    SUR_TRY(s, hipSetDevice(s->device));
    if (int rc = upload_x(s, n, x)) return rc;
    if (int rc = forward(s, n)) return rc;
    SUR_TRY(s, hipMemcpyAsync(out, s->Z[s->L - 1].p, (size_t)n * s->w[s->L] * 4, hipMemcpyDeviceToHost, s->stream));
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API int arbo_sur_forward_noalloc(ArboSurrogate* s, uint32_t n, const float* x, float* out, void* stream) {
    if (!s || !x || !out || n == 0) return ARBO_E_INVALID;
    uint32_t widest = 0;   // widest hidden layer: the activations held between layers
    for (uint32_t l = 1; l < s->L; ++l) widest = std::max(widest, s->w[l]);
    if ((uint64_t)n * widest > kScratchElems)
        return fail(s, ARBO_E_CAPACITY, "rows x widest hidden layer exceed the 2^20 activations of the no-allocation scratch");
    if (!s->scratch) return fail(s, ARBO_E_STATE, "no scratch address on this device");
    SUR_TRY(s, hipSetDevice(s->device));
    const hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    float* buf[2] = {s->scratch, s->scratch + kScratchElems};
    for (uint32_t l = 0; l < s->L; ++l) {   // the products of forward(), activations ping-ponged in the scratch
        const uint32_t K = s->w[l], N = s->w[l + 1];
        const bool last = l + 1 == s->L;
        GemmArgs g = args(n, N, K, last ? EPI_BIAS : EPI_GELU_ONLY);
        g.A = l == 0 ? x : buf[(l + 1) & 1];
        g.sAi = K;
        g.sAk = 1;
        g.B = s->P.p + s->woff[l];
        g.sBj = K;
        g.sBk = 1;
        g.bias = s->P.p + s->boff[l];
        if (last) {
            g.outf = out;
            g.ldf = N;
        } else {
            g.outh = buf[l & 1];
            g.ldh = N;
        }
        SUR_TRY(s, gemm(st, g));
    }
    return ARBO_OK;
}

ARBO_API int arbo_sur_input_grad(ArboSurrogate* s, uint32_t n, const float* x, const float* dout, float* out,
                                 float* dx) {
    if (!s || !x || !dout || !dx || n == 0) return ARBO_E_INVALID;
    if (n > s->max_batch) return fail(s, ARBO_E_CAPACITY, "more rows than max_batch");
    SUR_TRY(s, hipSetDevice(s->device));
    if (int rc = upload_x(s, n, x)) return rc;
    if (int rc = forward(s, n)) return rc;
    const uint64_t no = (uint64_t)n * s->w[s->L];
    SUR_TRY(s, hipMemcpyAsync(s->dZ[s->L - 1].p, dout, no * 4, hipMemcpyHostToDevice, s->stream));
    if (int rc = backward(s, n, false, true)) return rc;
    if (out) SUR_TRY(s, hipMemcpyAsync(out, s->Z[s->L - 1].p, no * 4, hipMemcpyDeviceToHost, s->stream));
    SUR_TRY(s, hipMemcpyAsync(dx, s->dxin.p, (size_t)n * s->w[0] * 4, hipMemcpyDeviceToHost, s->stream));
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API int arbo_sur_loss_grad(ArboSurrogate* s, uint32_t n, const float* x, const float* y, double* loss,
                                float* grad) {
    if (!s || !x || !y || n == 0) return ARBO_E_INVALID;
    if (n > s->max_batch) return fail(s, ARBO_E_CAPACITY, "more rows than max_batch");
    SUR_TRY(s, hipSetDevice(s->device));
    if (int rc = upload_x(s, n, x)) return rc;
    const uint64_t no = (uint64_t)n * s->w[s->L];
    SUR_TRY(s, hipMemcpyAsync(s->yin.p, y, no * 4, hipMemcpyHostToDevice, s->stream));
    if (int rc = forward(s, n)) return rc;
    double l = 0.0;
    if (int rc = poisson(s, n, s->yin.p, 1.0f / (float)no, &l)) return rc;
    if (loss) *loss = l / (double)no;
    if (int rc = backward(s, n, true, false)) return rc;
    if (grad) SUR_TRY(s, hipMemcpyAsync(grad, s->G.p, s->n_params * 4, hipMemcpyDeviceToHost, s->stream));
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API int arbo_sur_train(ArboSurrogate* s, uint32_t n, const float* x, const float* y, uint32_t epochs,
                            uint32_t batch, float lr, uint64_t seed, double* epoch_loss) {
    if (!s || !x || !y || n == 0 || batch == 0 || !(lr > 0.0f)) return ARBO_E_INVALID;
    if (batch > s->max_batch) return fail(s, ARBO_E_CAPACITY, "batch larger than max_batch");
    SUR_TRY(s, hipSetDevice(s->device));
    const uint32_t in = s->w[0], out = s->w[s->L];
    int rc = alloc(s, s->data_x, (size_t)n * in);
    if (!rc) rc = alloc(s, s->data_y, (size_t)n * out);
    if (!rc) rc = alloc(s, s->idx, n);
    if (rc) return rc;
    SUR_TRY(s, hipMemcpyAsync(s->data_x.p, x, (size_t)n * in * 4, hipMemcpyHostToDevice, s->stream));
    SUR_TRY(s, hipMemcpyAsync(s->data_y.p, y, (size_t)n * out * 4, hipMemcpyHostToDevice, s->stream));
    const Key key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    std::vector<uint32_t> perm(n);
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        for (uint32_t i = n - 1; i > 0; --i) {   // Fisher-Yates on Philox draws
            const U4 r = draw(key, i, e, 0x7u, 0);
            const uint32_t j = mulhi(r.x, i + 1);
            const uint32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        SUR_TRY(s, hipStreamSynchronize(s->stream));   // perm is reused below
        SUR_TRY(s, hipMemcpyAsync(s->idx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
        if (epoch_loss) {
            hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<float*>(s->epoch_sum.p),
                               (uint64_t)2);
            SUR_TRY(s, hipGetLastError());
        }
        for (uint32_t b0 = 0; b0 < n; b0 += batch) {
            const uint32_t B = std::min(batch, n - b0);
            hipLaunchKernelGGL(k_gather, dim3(blocks((uint64_t)B * in, 256)), dim3(256), 0, s->stream, s->data_x.p,
                               s->idx.p + b0, s->H[0].p, B, in);
            SUR_TRY(s, hipGetLastError());
            if ((rc = forward(s, B))) return rc;
            // each batch's loss before its update, summed on the device; the
            // targets are read through the batch's row indices
            if ((rc = poisson(s, B, s->data_y.p, 1.0f / ((float)B * out), nullptr,
                              epoch_loss ? s->epoch_sum.p : nullptr, s->idx.p + b0)))
                return rc;
            if ((rc = backward(s, B, true, false))) return rc;
            if ((rc = adam_step(s, lr))) return rc;
        }
        if (epoch_loss) {   // one sync per epoch
            double tot = 0.0;
            SUR_TRY(s, hipMemcpyAsync(&tot, s->epoch_sum.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
            SUR_TRY(s, hipStreamSynchronize(s->stream));
            epoch_loss[e] = tot / ((double)n * out);
        }
    }
    SUR_TRY(s, hipStreamSynchronize(s->stream));
    return ARBO_OK;
}

ARBO_API void arbo_sur_destroy(ArboSurrogate* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    release(s->P);
    release(s->G);
    release(s->Mo);
    release(s->Vo);
    for (auto& b : s->Z) release(b);
    for (auto& b : s->H) release(b);
    for (auto& b : s->dZ) release(b);
    release(s->yin);
    release(s->dxin);
    release(s->part);
    release(s->epoch_sum);
    release(s->data_x);
    release(s->data_y);
    release(s->idx);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

ARBO_API const char* arbo_sur_last_error(const ArboSurrogate* s) { return s ? s->err.c_str() : "null surrogate"; }

}  // extern "C"

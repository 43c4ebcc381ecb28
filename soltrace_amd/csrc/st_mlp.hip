// The flux surrogate's network: a multilayer perceptron with BF16 weights
// and activations and FP32 accumulation, trained and evaluated by the
// kernels here on the MFMA units of gfx950.
//
//   x (n_in)  -> [Linear + ReLU] x depth (width) -> Linear -> y (n_out)
//
// Layout in HBM (row-major, one row per sample):
//   X      bf16 [Mcap][64]      inputs, zero-padded to 64 columns
//   H[l]   bf16 [Mcap][width]   ReLU outputs of hidden layer l
//   Y      f32  [Mcap][n_out]   outputs
//   W[l]   f32 master, bf16 copy, Adam m and v, all [N][Kp] (PyTorch's
//          Linear layout: a row per output); biases f32 [N]
//   dZ     bf16 [Mcap][max(width, n_out)], two of them: the gradient with
//          respect to the current layer's pre-activation
//   gW[l]  f32 [splits][N][Kp] split-K partial weight gradients, summed by
//          the Adam kernel; gb[l] f32 [row blocks][N] bias partials
//
// One GEMM kernel serves the three products of a step:
//   forward   Y = X W^T          A = X [M][K],      B = W [N][K]
//   dX        dX = dZ W          A = dZ [M][Kred],  B = W stored [Kred][N]
//   dW        dW = dZ^T X        A = dZ stored [Kred][M], B = X stored [Kred][N]
// Operand tiles are staged through LDS in a [row][k] layout whatever their
// layout in HBM, so the MFMA fragment loads are one 16-byte LDS read per
// lane; the epilogue adds the bias, applies ReLU or its mask and converts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#define STMLP_API extern "C" __attribute__((visibility("default")))

namespace st {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Round to nearest even (inputs here are finite).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

constexpr int kTM = 64, kTN = 64, kTK = 32, kThreads = 256;
constexpr int kLds = kTK + 8;  // bf16 per LDS row: 80 B, so the 16 rows a quarter-wave reads start on distinct banks
constexpr int kInPad = 64;     // input columns after padding (one N tile of the first layer's dW)
constexpr int kMaxSeg = 32;    // tensors one Adam launch updates

enum Epi : int {
    EPI_F32 = 0,             // C (f32, split z at C + z * c_split) = acc
    EPI_BIAS_RELU_BF16 = 1,  // C (bf16) = relu(acc + bias)
    EPI_BIAS_F32 = 2,        // C (f32) = acc + bias, rows < m_store
    EPI_MASK_BF16 = 3,       // C (bf16) = mask > 0 ? acc : 0 (ReLU backward)
};

struct GemmArgs {
    const uint16_t *A;
    long lda;
    const uint16_t *B;
    long ldb;
    int M, N, K;  // K: reduction length of the whole launch
    int kchunk;   // reduction length of one blockIdx.z
    const float *bias;
    const uint16_t *mask;
    long ldm;
    void *C;
    long ldc;
    long c_split;
    int m_store;
};

// One 64 x 32 operand tile: a 16-byte global load per thread.
//   T = false: P[r][k] (k contiguous): thread -> row tid/4, k chunk tid%4
//   T = true:  P[k][r] (r contiguous): thread -> k tid/8, row chunk tid%8
template <bool T>
__device__ __forceinline__ uint4 load_tile(const uint16_t *P, long ld, int r0, int k0, int tid) {
    if (!T) return *(const uint4 *)(P + (long)(r0 + (tid >> 2)) * ld + k0 + (tid & 3) * 8);
    return *(const uint4 *)(P + (long)(k0 + (tid >> 3)) * ld + r0 + (tid & 7) * 8);
}
template <bool T>
__device__ __forceinline__ void store_tile(uint16_t (*S)[kLds], int tid, const uint4 &v) {
    if (!T) {
        *(uint4 *)&S[tid >> 2][(tid & 3) * 8] = v;
    } else {
        const int k = tid >> 3, r = (tid & 7) * 8;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            S[r + 2 * i][k] = (uint16_t)(w[i] & 0xffffu);
            S[r + 2 * i + 1][k] = (uint16_t)(w[i] >> 16);
        }
    }
}

// C[M][N] = A[M][K] B[K][N] over this block's 64 x 64 tile and blockIdx.z's
// K range; four waves, each a 32 x 32 MFMA accumulator
// (v_mfma_f32_32x32x16_bf16: lane l holds row l%32 of A and column l%32 of
// B at k = 8*(l/32) .. +7; accumulator i of lane l is row
// 8*(i/4) + 4*(l/32) + i%4, column l%32).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(kThreads) void gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kTM][kLds];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kTN][kLds];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
    uint4 va = load_tile<TA>(g.A, g.lda, m0, kb, tid);
    uint4 vb = load_tile<TB>(g.B, g.ldb, n0, kb, tid);
    const int ar = wm * 32 + (lane & 31), bc = wn * 32 + (lane & 31), kq = (lane >> 5) * 8;
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        __syncthreads();
        store_tile<TA>(As, tid, va);
        store_tile<TB>(Bs, tid, vb);
        __syncthreads();
        if (k0 + kTK < ke) {
            va = load_tile<TA>(g.A, g.lda, m0, k0 + kTK, tid);
            vb = load_tile<TB>(g.B, g.ldb, n0, k0 + kTK, tid);
        }
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)&As[ar][kk + kq]);
            const bf16x8 b = __builtin_bit_cast(bf16x8, *(const uint4 *)&Bs[bc][kk + kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    const int col = n0 + bc;
    const float bias = (EPI == EPI_BIAS_RELU_BF16 || EPI == EPI_BIAS_F32) ? g.bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int row = m0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        float v = acc[i];
        if constexpr (EPI == EPI_F32) {
            ((float *)g.C)[(long)blockIdx.z * g.c_split + (long)row * g.ldc + col] = v;
        } else if constexpr (EPI == EPI_BIAS_RELU_BF16) {
            v = fmaxf(v + bias, 0.0f);
            ((uint16_t *)g.C)[(long)row * g.ldc + col] = f2bf(v);
        } else if constexpr (EPI == EPI_BIAS_F32) {
            if (row < g.m_store) ((float *)g.C)[(long)row * g.ldc + col] = v + bias;
        } else {
            const uint16_t h = g.mask[(long)row * g.ldm + col];
            const bool on = h != 0 && !(h & 0x8000u);
            ((uint16_t *)g.C)[(long)row * g.ldc + col] = on ? f2bf(v) : (uint16_t)0;
        }
    }
}

// x (f32 [b][n_in]) -> X (bf16 [M][kInPad]), zero padding in both directions.
__global__ void pack_input_kernel(const float *x, int b, int n_in, int M, uint16_t *X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * kInPad) return;
    const int r = i / kInPad, c = i % kInPad;
    X[i] = (r < b && c < n_in) ? f2bf(x[(long)r * n_in + c]) : (uint16_t)0;
}

// Mean squared error over the b real rows: dY = 2 (Y - T) / (b n), bf16;
// part[block] = sum of squared differences of the block's elements.  Rows
// b..M-1 get dY = 0, so padding contributes nothing to any gradient.
constexpr int kLossBlock = 256;
__global__ __launch_bounds__(kLossBlock) void mse_kernel(const float *Y, const uint16_t *T, int b, int n, int M,
                                                        float scale, uint16_t *dY, float *part) {
    __shared__ float red[kLossBlock];
    const long total = (long)M * n, real = (long)b * n;
    float s = 0.0f;
    for (long i = (long)blockIdx.x * kLossBlock + threadIdx.x; i < total; i += (long)gridDim.x * kLossBlock) {
        if (i < real) {
            const float d = Y[i] - bf2f(T[i]);
            s += d * d;
            dY[i] = f2bf(scale * d);
        } else {
            dY[i] = 0;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kLossBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Bias gradient partials: out[s][n] = sum of dZ[r][n] over row block s
// (16 rows, so a batch of 512 gives 32 blocks of rows per column block).
constexpr int kColRows = 16;
__global__ void colsum_kernel(const uint16_t *dZ, long ld, int M, int N, float *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.y * kColRows, r1 = min(M, r0 + kColRows);
    float s = 0.0f;
    for (int r = r0; r < r1; r++) s += bf2f(dZ[(long)r * ld + n]);
    out[(long)blockIdx.y * N + n] = s;
}

// Split-K partials of a masked product (f32 [splits][M][N], c_split apart)
// summed in split order, the ReLU mask applied and converted, as the
// EPI_MASK_BF16 epilogue does for one split.
__global__ void mask_sum_kernel(const float *part, int splits, long c_split, int M, int N, const uint16_t *mask,
                                long ldm, uint16_t *C, long ldc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * N) return;
    const int r = (int)(i / N), c = (int)(i % N);
    float v = 0.0f;
    for (int k = 0; k < splits; k++) v += part[(long)k * c_split + (long)r * N + c];
    const uint16_t h = mask[(long)r * ldm + c];
    const bool on = h != 0 && !(h & 0x8000u);
    C[(long)r * ldc + c] = on ? f2bf(v) : (uint16_t)0;
}

// One tensor of an Adam launch: parameters, their bf16 copy, moments and
// the split partials of the gradient (g + s * gstride, s < nsplit).
struct Seg {
    float *p, *m, *v;
    uint16_t *pb;
    const float *g;
    long n, gstride;
    int nsplit;
    int first_block;
};
struct AdamArgs {
    Seg seg[kMaxSeg];
    int nseg;
    float lr, b1, b2, eps, bc1, bc2;
    // graph replays: lr, bc1, bc2 of step *step_ctr from table (3 floats a step)
    const float *table;
    const int *step_ctr;
    int update;   // 0: only sum the partials into grad_out
    float *grad_out;  // (may be null) summed gradient, segments laid end to end
    long grad_off[kMaxSeg];
};
__global__ void adam_kernel(AdamArgs a) {
    int s = 0;
    while (s + 1 < a.nseg && (int)blockIdx.x >= a.seg[s + 1].first_block) s++;
    const Seg &q = a.seg[s];
    const long i = (long)(blockIdx.x - q.first_block) * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    float g = 0.0f;
    for (int k = 0; k < q.nsplit; k++) g += q.g[(long)k * q.gstride + i];
    if (a.grad_out) a.grad_out[a.grad_off[s] + i] = g;
    if (!a.update) return;
    float lr = a.lr, bc1 = a.bc1, bc2 = a.bc2;
    if (a.table) {
        const float *t = a.table + 3 * (long)*a.step_ctr;
        lr = t[0];
        bc1 = t[1];
        bc2 = t[2];
    }
    const float m = a.b1 * q.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * q.v[i] + (1.0f - a.b2) * g * g;
    q.m[i] = m;
    q.v[i] = v;
    const float p = q.p[i] - lr * (m / bc1) / (sqrtf(v / bc2) + a.eps);
    q.p[i] = p;
    if (q.pb) q.pb[i] = f2bf(p);  // biases have no bf16 copy: the epilogue reads the f32 value
}

__global__ void to_bf16_kernel(const float *p, uint16_t *pb, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

// The batch of step *step_ctr of a resident training set: rows
// idx[step * b + r] of x (f32 [rows][n_in]) packed as pack_input_kernel
// packs them, and of t (bf16 [rows][n_out]) into tb; padding rows of tb zero.
// An index outside [0, rows) is clamped into it.
__global__ void gather_batch_kernel(const float *x, const uint16_t *t, const int *idx, const int *step_ctr, int b,
                                    long rows, int n_in, int n_out, int M, uint16_t *X, uint16_t *tb) {
    const long per_row = kInPad + n_out;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * per_row) return;
    const int r = (int)(i / per_row), c = (int)(i % per_row);
    long src = -1;
    if (r < b) {
        src = (long)idx[(long)*step_ctr * b + r];
        src = src < 0 ? 0 : (src >= rows ? rows - 1 : src);
    }
    if (c < kInPad) X[(long)r * kInPad + c] = (src >= 0 && c < n_in) ? f2bf(x[src * n_in + c]) : (uint16_t)0;
    else tb[(long)r * n_out + (c - kInPad)] = src >= 0 ? t[src * n_out + (c - kInPad)] : (uint16_t)0;
}
__global__ void set_int_kernel(int *p, int v) { *p = v; }
__global__ void bump_kernel(int *p) { *p += 1; }

// ------------------------------------------------------------- host -----
inline void hc(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        char b[256];
        snprintf(b, sizeof b, "HIP error in %s: %s", what, hipGetErrorString(e));
        throw std::runtime_error(b);
    }
}
#define MC(x) hc((x), #x)

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    template <typename T> T *as() const { return (T *)p; }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    void alloc(size_t bytes) {
        MC(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        MC(hipMemset(p, 0, std::max<size_t>(bytes, 16)));
    }
};

struct Layer {
    int K = 0, Kp = 0, N = 0;  // inputs (real, padded), outputs
    int wsplit = 1;            // split-K parts of the weight gradient buffer
    int wsplit_used = 1;       // parts the last backward pass wrote
    Buf W, Wb, mW, vW, b, mb, vb, gW, gb;
};

struct Net {
    int n_in, width, depth, n_out, max_batch, Mcap, rowblocks;
    std::vector<Layer> L;
    Buf X, Y, dZ[2], loss_part;
    Buf dXpart;          // split-K partials of the widest layer's input gradient
    int dx_splits = 1;
    std::vector<Buf> H;  // depth hidden activations
    int rowblocks_used = 1;  // bias partials the last backward pass wrote
    int step = 0;
    int lossblocks = 256;
    std::vector<float> part_host;
    // graph replays (fit): the gathered targets, the step counter, the
    // per-step table, and a private stream for capture
    Buf Tb, ctr, table;
    size_t table_steps = 0;
    hipStream_t cap = nullptr;
    ~Net() {
        if (cap) (void)hipStreamDestroy(cap);
    }
};

thread_local std::string g_error;

long long param_count(const Net &n) {
    long long c = 0;
    for (const Layer &l : n.L) c += (long long)l.N * l.K + l.N;
    return c;
}

template <bool TA, bool TB, int EPI>
void gemm(const GemmArgs &g, int splits, hipStream_t s) {
    if (g.M % kTM || g.N % kTN || g.kchunk % kTK || g.lda % 8 || g.ldb % 8)
        throw std::runtime_error("gemm: shape not a multiple of the tile");
    dim3 grid(g.N / kTN, g.M / kTM, splits);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(kThreads), 0, s, g);
    MC(hipGetLastError());
}

int rows_padded(int b) { return (b + kTM - 1) / kTM * kTM; }

// Forward pass over b rows of x (device f32 [b][n_in]); the last layer
// writes f32 rows < b to y (ldc n_out) when y is given, else to net.Y.
void forward(Net &n, const float *x, int b, float *y, hipStream_t s) {
    const int M = rows_padded(b);
    if (x) {  // x == nullptr: net.X already holds the packed batch
        const int tot = M * kInPad;
        hipLaunchKernelGGL(pack_input_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, x, b, n.n_in, M,
                           n.X.as<uint16_t>());
        MC(hipGetLastError());
    }
    const uint16_t *in = n.X.as<uint16_t>();
    long ldin = kInPad;
    for (int l = 0; l <= n.depth; l++) {
        const Layer &ly = n.L[l];
        GemmArgs g{};
        g.A = in;
        g.lda = ldin;
        g.B = ly.Wb.as<uint16_t>();
        g.ldb = ly.Kp;
        g.M = M;
        g.N = ly.N;
        g.K = ly.Kp;
        g.kchunk = ly.Kp;
        g.bias = ly.b.as<float>();
        if (l < n.depth) {
            g.C = n.H[l].p;
            g.ldc = n.width;
            gemm<false, false, EPI_BIAS_RELU_BF16>(g, 1, s);
            in = n.H[l].as<uint16_t>();
            ldin = n.width;
        } else {
            g.C = y ? (void *)y : n.Y.p;
            g.ldc = n.n_out;
            g.m_store = y ? b : M;
            gemm<false, false, EPI_BIAS_F32>(g, 1, s);
        }
    }
}

// Loss and backward pass after forward(x, b) into net.Y; leaves the split
// partials of every gradient in gW / gb.
void backward(Net &n, const uint16_t *t, int b, hipStream_t s) {
    const int M = rows_padded(b);
    const int D = n.n_out;
    const float scale = 2.0f / ((float)b * (float)D);
    hipLaunchKernelGGL(mse_kernel, dim3(n.lossblocks), dim3(kLossBlock), 0, s, n.Y.as<float>(), t, b, D, M, scale,
                       n.dZ[0].as<uint16_t>(), n.loss_part.as<float>());
    MC(hipGetLastError());
    int cur = 0;
    for (int l = n.depth; l >= 0; l--) {
        Layer &ly = n.L[l];
        const uint16_t *dz = n.dZ[cur].as<uint16_t>();
        const long lddz = ly.N;
        const uint16_t *hin = l == 0 ? n.X.as<uint16_t>() : n.H[l - 1].as<uint16_t>();
        const long ldh = l == 0 ? kInPad : n.width;
        // dW[N][Kp] = dZ^T H: A = dZ stored [M][N] (transposed), B = H stored [M][Kp] (transposed)
        {
            GemmArgs g{};
            g.A = dz;
            g.lda = lddz;
            g.B = hin;
            g.ldb = ldh;
            g.M = ly.N;
            g.N = ly.Kp;
            g.K = M;
            int splits = std::max(1, std::min(ly.wsplit, M / kTK));
            int chunk = (M / splits + kTK - 1) / kTK * kTK;
            splits = (M + chunk - 1) / chunk;
            g.kchunk = chunk;
            g.C = ly.gW.p;
            g.ldc = ly.Kp;
            g.c_split = (long)ly.N * ly.Kp;
            gemm<true, true, EPI_F32>(g, splits, s);
            ly.wsplit_used = splits;  // the Adam pass sums exactly these
        }
        {
            const int used = (M + kColRows - 1) / kColRows;
            dim3 grid((ly.N + 255) / 256, used);
            hipLaunchKernelGGL(colsum_kernel, grid, dim3(256), 0, s, dz, lddz, M, ly.N, ly.gb.as<float>());
            MC(hipGetLastError());
            n.rowblocks_used = used;
        }
        if (l > 0) {
            // dH[M][K] = dZ W, masked by H > 0: A = dZ [M][N], B = W stored [N][Kp] (transposed).
            // A long reduction (the output layer's) is split over blockIdx.z
            // into f32 partials, summed and masked by a second pass.
            GemmArgs g{};
            g.A = dz;
            g.lda = lddz;
            g.B = ly.Wb.as<uint16_t>();
            g.ldb = ly.Kp;
            g.M = M;
            g.N = ly.Kp;
            g.K = ly.N;
            const int splits = (ly.N >= 1024 && n.dx_splits > 1) ? n.dx_splits : 1;
            g.kchunk = ly.N / splits;
            if (splits > 1) {
                g.C = n.dXpart.p;
                g.ldc = ly.Kp;
                g.c_split = (long)M * ly.Kp;
                gemm<false, true, EPI_F32>(g, splits, s);
                const long tot = (long)M * ly.Kp;
                hipLaunchKernelGGL(mask_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                                   n.dXpart.as<float>(), splits, g.c_split, M, ly.Kp, hin, ldh,
                                   n.dZ[cur ^ 1].as<uint16_t>(), (long)ly.Kp);
                MC(hipGetLastError());
            } else {
                g.kchunk = ly.N;
                g.mask = hin;
                g.ldm = ldh;
                g.C = n.dZ[cur ^ 1].p;
                g.ldc = ly.Kp;
                gemm<false, true, EPI_MASK_BF16>(g, 1, s);
            }
            cur ^= 1;
        }
    }
}

// Sum the gradient partials; with update, take an Adam step.
void adam(Net &n, bool update, float lr, float *grad_out, hipStream_t s, const float *table = nullptr,
          const int *step_ctr = nullptr) {
    AdamArgs a{};
    a.table = table;
    a.step_ctr = step_ctr;
    int blocks = 0;
    long off = 0;
    auto add = [&](float *p, float *m, float *v, uint16_t *pb, const float *g, long cnt, long gstride, int nsplit) {
        if (a.nseg >= kMaxSeg) throw std::runtime_error("too many tensors");
        Seg &q = a.seg[a.nseg];
        q.p = p;
        q.m = m;
        q.v = v;
        q.pb = pb;
        q.g = g;
        q.n = cnt;
        q.gstride = gstride;
        q.nsplit = nsplit;
        q.first_block = blocks;
        a.grad_off[a.nseg] = off;
        a.nseg++;
        blocks += (int)((cnt + 255) / 256);
        off += cnt;
    };
    for (Layer &l : n.L) {
        const long nw = (long)l.N * l.Kp;
        add(l.W.as<float>(), l.mW.as<float>(), l.vW.as<float>(), l.Wb.as<uint16_t>(), l.gW.as<float>(), nw, nw,
            l.wsplit_used);
        add(l.b.as<float>(), l.mb.as<float>(), l.vb.as<float>(), nullptr, l.gb.as<float>(), l.N, l.N,
            n.rowblocks_used);
    }
    a.update = update ? 1 : 0;
    a.grad_out = grad_out;
    if (update) {
        n.step++;
        a.lr = lr;
        a.b1 = 0.9f;
        a.b2 = 0.999f;
        a.eps = 1e-8f;
        a.bc1 = 1.0f - powf(a.b1, (float)n.step);
        a.bc2 = 1.0f - powf(a.b2, (float)n.step);
    }
    hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, s, a);
    MC(hipGetLastError());
}

// splitmix64: the initial weights depend only on the seed
struct Mix {
    uint64_t x;
    uint64_t next() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

Net *create(int n_in, int width, int depth, int n_out, int max_batch, unsigned long long seed) {
    if (n_in < 1 || n_in > kInPad || width < kTN || width % kTN || depth < 1 || depth > 14 || n_out < kTN ||
        n_out % kTN || max_batch < 1)
        throw std::runtime_error("mlp: unsupported shape (n_in <= 64, width and n_out multiples of 64, "
                                 "1 <= depth <= 14)");
    std::unique_ptr<Net> n(new Net());
    n->n_in = n_in;
    n->width = width;
    n->depth = depth;
    n->n_out = n_out;
    n->max_batch = max_batch;
    n->Mcap = rows_padded(max_batch);
    n->rowblocks = (n->Mcap + kColRows - 1) / kColRows;
    n->L.resize(depth + 1);
    Mix rng{seed};
    for (int l = 0; l <= depth; l++) {
        Layer &ly = n->L[l];
        ly.K = l == 0 ? n_in : width;
        ly.Kp = l == 0 ? kInPad : width;
        ly.N = l == depth ? n_out : width;
        // split-K parts of the weight gradient: enough blocks to fill the chip
        const int tiles = (ly.N / kTM) * (ly.Kp / kTN);
        ly.wsplit = std::max(1, std::min(n->Mcap / kTK, (1024 + tiles - 1) / tiles));
        const long nw = (long)ly.N * ly.Kp;
        for (Buf *b : {&ly.W, &ly.mW, &ly.vW}) b->alloc(nw * sizeof(float));
        ly.Wb.alloc(nw * sizeof(uint16_t));
        for (Buf *b : {&ly.b, &ly.mb, &ly.vb}) b->alloc(ly.N * sizeof(float));
        ly.gW.alloc((size_t)ly.wsplit * nw * sizeof(float));
        ly.gb.alloc((size_t)n->rowblocks * ly.N * sizeof(float));
        // He-uniform weights for the ReLU layers, a smaller range for the
        // linear output layer; zero biases; padded input columns zero
        std::vector<float> w(nw, 0.0f);
        const double a = l < depth ? std::sqrt(6.0 / ly.K) : std::sqrt(1.0 / ly.K);
        for (int i = 0; i < ly.N; i++)
            for (int k = 0; k < ly.K; k++) w[(long)i * ly.Kp + k] = (float)((2.0 * rng.uniform() - 1.0) * a);
        MC(hipMemcpy(ly.W.p, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_bf16_kernel, dim3((nw + 255) / 256), dim3(256), 0, 0, ly.W.as<float>(),
                           ly.Wb.as<uint16_t>(), nw);
        MC(hipGetLastError());
    }
    n->X.alloc((size_t)n->Mcap * kInPad * sizeof(uint16_t));
    n->H.resize(depth);
    for (Buf &h : n->H) h.alloc((size_t)n->Mcap * width * sizeof(uint16_t));
    n->Y.alloc((size_t)n->Mcap * n_out * sizeof(float));
    for (Buf &d : n->dZ) d.alloc((size_t)n->Mcap * std::max(width, n_out) * sizeof(uint16_t));
    n->loss_part.alloc(n->lossblocks * sizeof(float));
    // split-K of the output layer's input gradient: eight parts of n_out / 8
    // (a multiple of the 32-deep reduction step) when n_out allows it
    if (n_out >= 1024 && (n_out / 8) % kTK == 0) {
        n->dx_splits = 8;
        n->dXpart.alloc((size_t)8 * n->Mcap * width * sizeof(float));
    }
    n->part_host.resize(n->lossblocks);
    MC(hipDeviceSynchronize());
    return n.release();
}

double read_loss(Net &n, int b, hipStream_t s) {
    MC(hipMemcpyAsync(n.part_host.data(), n.loss_part.p, n.lossblocks * sizeof(float), hipMemcpyDeviceToHost, s));
    MC(hipStreamSynchronize(s));
    double t = 0.0;
    for (float v : n.part_host) t += v;
    return t / ((double)b * n.n_out);
}

// Parameters in the exchange layout: per layer W [N][K] (real inputs
// only), then b [N].  which 0: f32 master values, 1: the bf16 copies the
// kernels compute with (as f32).
void get_params(Net &n, float *out, int which) {
    long o = 0;
    for (Layer &l : n.L) {
        const long nw = (long)l.N * l.Kp;
        std::vector<float> w(nw);
        if (which == 1) {
            std::vector<uint16_t> wb(nw);
            MC(hipMemcpy(wb.data(), l.Wb.p, nw * sizeof(uint16_t), hipMemcpyDeviceToHost));
            for (long i = 0; i < nw; i++) w[i] = bf2f(wb[i]);
        } else {
            MC(hipMemcpy(w.data(), l.W.p, nw * sizeof(float), hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < l.N; i++)
            for (int k = 0; k < l.K; k++) out[o++] = w[(long)i * l.Kp + k];
        MC(hipMemcpy(out + o, l.b.p, l.N * sizeof(float), hipMemcpyDeviceToHost));
        o += l.N;
    }
}

void set_params(Net &n, const float *in) {
    long o = 0;
    MC(hipDeviceSynchronize());
    for (Layer &l : n.L) {
        const long nw = (long)l.N * l.Kp;
        std::vector<float> w(nw, 0.0f), zero(nw, 0.0f);
        for (int i = 0; i < l.N; i++)
            for (int k = 0; k < l.K; k++) w[(long)i * l.Kp + k] = in[o++];
        MC(hipMemcpy(l.W.p, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        MC(hipMemcpy(l.mW.p, zero.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        MC(hipMemcpy(l.vW.p, zero.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        MC(hipMemcpy(l.b.p, in + o, l.N * sizeof(float), hipMemcpyHostToDevice));
        MC(hipMemcpy(l.mb.p, zero.data(), l.N * sizeof(float), hipMemcpyHostToDevice));
        MC(hipMemcpy(l.vb.p, zero.data(), l.N * sizeof(float), hipMemcpyHostToDevice));
        o += l.N;
        hipLaunchKernelGGL(to_bf16_kernel, dim3((nw + 255) / 256), dim3(256), 0, 0, l.W.as<float>(),
                           l.Wb.as<uint16_t>(), nw);
        MC(hipGetLastError());
    }
    n.step = 0;
    MC(hipDeviceSynchronize());
}

// Gradient in the exchange layout from the summed (padded) gradient.
void unpad_grad(const Net &n, const std::vector<float> &g, float *out) {
    long o = 0, src = 0;
    for (const Layer &l : n.L) {
        for (int i = 0; i < l.N; i++)
            for (int k = 0; k < l.K; k++) out[o++] = g[src + (long)i * l.Kp + k];
        src += (long)l.N * l.Kp;
        for (int i = 0; i < l.N; i++) out[o++] = g[src + i];
        src += l.N;
    }
}

long padded_count(const Net &n) {
    long c = 0;
    for (const Layer &l : n.L) c += (long)l.N * l.Kp + l.N;
    return c;
}

void check_batch(const Net &n, int b) {
    if (b < 1 || b > n.max_batch) throw std::runtime_error("mlp: batch outside 1..max_batch");
}

// `steps` Adam steps on batches of a resident training set: one step
// (gather, forward, loss, backward, Adam, counter) is captured into a HIP
// graph on a private stream and replayed `steps` times on s.  The step's
// batch, learning rate and bias corrections are read on the device at the
// step counter, so every replay is the same graph.  The bias corrections
// are computed here as the eager step computes them, so the two paths give
// the same bits.
double fit(Net &n, const float *x, const uint16_t *t, long rows, const int *idx, int b, int steps, const float *lr,
           bool want_loss, hipStream_t s) {
    if (steps < 1 || rows < 1) throw std::runtime_error("mlp: fit needs steps >= 1 and rows >= 1");
    check_batch(n, b);
    const int M = rows_padded(b);
    if (!n.Tb.p) n.Tb.alloc((size_t)n.Mcap * n.n_out * sizeof(uint16_t));
    if (!n.ctr.p) n.ctr.alloc(sizeof(int));
    if (n.table_steps < (size_t)steps) {
        n.table.release();
        n.table.alloc((size_t)steps * 3 * sizeof(float));
        n.table_steps = steps;
    }
    if (!n.cap) MC(hipStreamCreateWithFlags(&n.cap, hipStreamNonBlocking));
    // per-step lr and bias corrections (the host's powf, as adam() uses)
    std::vector<float> lr_h(steps), tab(3 * (size_t)steps);
    MC(hipMemcpyAsync(lr_h.data(), lr, steps * sizeof(float), hipMemcpyDeviceToHost, s));
    MC(hipStreamSynchronize(s));
    for (int k = 0; k < steps; k++) {
        const int st = n.step + 1 + k;
        tab[3 * k] = lr_h[k];
        tab[3 * k + 1] = 1.0f - powf(0.9f, (float)st);
        tab[3 * k + 2] = 1.0f - powf(0.999f, (float)st);
    }
    MC(hipMemcpyAsync(n.table.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, n.ctr.as<int>(), 0);
    MC(hipGetLastError());
    // the private stream starts after everything queued on s so far
    hipEvent_t ev;
    MC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    MC(hipEventRecord(ev, s));
    MC(hipStreamWaitEvent(n.cap, ev, 0));
    MC(hipStreamSynchronize(n.cap));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipEventDestroy(ev);
    };
    try {
        MC(hipStreamBeginCapture(n.cap, hipStreamCaptureModeThreadLocal));
        try {
            const long tot = (long)M * (kInPad + n.n_out);
            hipLaunchKernelGGL(gather_batch_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, n.cap, x, t,
                               idx, n.ctr.as<int>(), b, rows, n.n_in, n.n_out, M, n.X.as<uint16_t>(),
                               n.Tb.as<uint16_t>());
            MC(hipGetLastError());
            forward(n, nullptr, b, nullptr, n.cap);
            backward(n, n.Tb.as<uint16_t>(), b, n.cap);
            adam(n, true, 0.0f, nullptr, n.cap, n.table.as<float>(), n.ctr.as<int>());
            hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(1), 0, n.cap, n.ctr.as<int>());
            MC(hipGetLastError());
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(n.cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        MC(hipStreamEndCapture(n.cap, &graph));
        MC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (int k = 0; k < steps; k++) MC(hipGraphLaunch(exec, s));
        n.step += steps;  // adam() counted one step at capture
        n.step -= 1;
        double loss = want_loss ? read_loss(n, b, s) : 0.0;
        MC(hipStreamSynchronize(s));
        cleanup();
        return loss;
    } catch (...) {
        cleanup();
        throw;
    }
}

}  // namespace mlp
}  // namespace st

using namespace st::mlp;

namespace {
template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}
hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace

STMLP_API void *st_amd_mlp_create(int n_in, int width, int depth, int n_out, int max_batch,
                                  unsigned long long seed) {
    try {
        return create(n_in, width, depth, n_out, max_batch, seed);
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}
STMLP_API void st_amd_mlp_free(void *net) { delete reinterpret_cast<Net *>(net); }
STMLP_API const char *st_amd_mlp_error(void) { return g_error.c_str(); }
STMLP_API long long st_amd_mlp_num_params(void *net) {
    return net ? param_count(*reinterpret_cast<Net *>(net)) : -1;
}
STMLP_API int st_amd_mlp_get_params(void *net, float *out, int which) {
    return guarded([&] {
        if (!net || !out || (which != 0 && which != 1)) return -1;
        get_params(*reinterpret_cast<Net *>(net), out, which);
        return 1;
    });
}
STMLP_API int st_amd_mlp_set_params(void *net, const float *in) {
    return guarded([&] {
        if (!net || !in) return -1;
        set_params(*reinterpret_cast<Net *>(net), in);
        return 1;
    });
}
STMLP_API int st_amd_mlp_forward(void *net, const float *x, int batch, float *y, void *stream) {
    return guarded([&] {
        if (!net || !x || !y) return -1;
        Net &n = *reinterpret_cast<Net *>(net);
        check_batch(n, batch);
        forward(n, x, batch, y, as_stream(stream));
        return 1;
    });
}
STMLP_API int st_amd_mlp_train_step(void *net, const float *x, const unsigned short *t, int batch, float lr,
                                    double *loss, void *stream) {
    return guarded([&] {
        if (!net || !x || !t || !(lr >= 0.0f)) return -1;
        Net &n = *reinterpret_cast<Net *>(net);
        check_batch(n, batch);
        const hipStream_t s = as_stream(stream);
        forward(n, x, batch, nullptr, s);
        backward(n, t, batch, s);
        if (loss) *loss = read_loss(n, batch, s);  // the loss of the weights before this step
        adam(n, true, lr, nullptr, s);
        return 1;
    });
}
STMLP_API int st_amd_mlp_gradients(void *net, const float *x, const unsigned short *t, int batch, float *grad,
                                   double *loss, void *stream) {
    return guarded([&] {
        if (!net || !x || !t || !grad) return -1;
        Net &n = *reinterpret_cast<Net *>(net);
        check_batch(n, batch);
        const hipStream_t s = as_stream(stream);
        forward(n, x, batch, nullptr, s);
        backward(n, t, batch, s);
        const long np = padded_count(n);
        float *gsum = nullptr;
        MC(hipMallocAsync((void **)&gsum, np * sizeof(float), s));
        adam(n, false, 0.0f, gsum, s);
        std::vector<float> g(np);
        MC(hipMemcpyAsync(g.data(), gsum, np * sizeof(float), hipMemcpyDeviceToHost, s));
        MC(hipFreeAsync(gsum, s));
        if (loss) *loss = read_loss(n, batch, s);
        MC(hipStreamSynchronize(s));
        unpad_grad(n, g, grad);
        return 1;
    });
}
STMLP_API int st_amd_mlp_fit(void *net, const float *x, const unsigned short *t, long long rows, const int *idx,
                             int batch, int steps, const float *lr, double *loss, void *stream) {
    return guarded([&] {
        if (!net || !x || !t || !idx || !lr) return -1;
        Net &n = *reinterpret_cast<Net *>(net);
        const double l = fit(n, x, t, (long)rows, idx, batch, steps, lr, loss != nullptr, as_stream(stream));
        if (loss) *loss = l;
        return 1;
    });
}

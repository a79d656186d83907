// Message-passing network that predicts partial-match table sizes.
// No reference counterpart: cuTS orders queries only with the greedy score
// of src/graph.cpp:62-93 and never looks at how large its trie levels get
// (the sizes live in `lengths`, src/host_funcs.cu:95,153, and are dropped).
//
// Model (one sample = a query graph with <= 16 vertices, a prefix set S of
// its vertices, and per-vertex input features that include data-graph
// statistics):
//   h0 = x                                                   (24 features)
//   h_{l+1}[v] = ReLU(b_l + Ws_l^T h_l[v] + Wn_l^T sum_{u in N(v)} h_l[u]),  l = 0,1,2 (hidden 64)
//   c = [ sum_{v in S} h3[v] ; mean_v h3[v] ]                 (128)
//   y = bo + wo . ReLU(br + Wr^T c)
// and is trained on y ~ log(1 + |T(S)|) / 10 with a mean squared error.
//
// gfx950 mapping: one wave64 per sample, lane j = hidden unit j.  Weights are
// BF16 (kept twice: input-major for the forward pass, where lane j reads
// W[i][j] for a fixed i in one coalesced 128-byte load, and transposed for
// the backward pass); activations and every accumulation are FP32, with an
// FP32 master copy updated by Adam.  Gradients are summed per block in that
// block's own row of a partial-sum array, one wave per block so that lane j
// owns column j of every gradient matrix (no atomics), then reduced over
// blocks.  The network is a chain of
// 16x64x64 products per sample -- far below MFMA tile sizes and bound by
// latency, not by any roofline.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "mpnn.hpp"

namespace cuts {
namespace {

constexpr int H = kMpnnHidden, FIN = kMpnnFeats, MAXN = kMpnnNodes;

// parameter offsets (floats) in the flat vector; matrices are input-major
// W[i][j] = W[i*out + j]
constexpr int O_WS0 = 0;
constexpr int O_WN0 = O_WS0 + FIN * H;
constexpr int O_B0 = O_WN0 + FIN * H;
constexpr int O_WS1 = O_B0 + H;
constexpr int O_WN1 = O_WS1 + H * H;
constexpr int O_B1 = O_WN1 + H * H;
constexpr int O_WS2 = O_B1 + H;
constexpr int O_WN2 = O_WS2 + H * H;
constexpr int O_B2 = O_WN2 + H * H;
constexpr int O_WR = O_B2 + H;
constexpr int O_BR = O_WR + 2 * H * H;
constexpr int O_WO = O_BR + H;
constexpr int O_BO = O_WO + H;
constexpr int NPARAM = O_BO + 1;
static_assert(NPARAM == kMpnnParams, "parameter layout");

__device__ __forceinline__ float bf(const uint16_t* w, int i) { return __uint_as_float((uint32_t)w[i] << 16); }

// round-to-nearest-even float -> bf16 (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Acts {
  float h0[MAXN][FIN], a0[MAXN][FIN];
  float h1[MAXN][H], a1[MAXN][H];
  float h2[MAXN][H], a2[MAXN][H];
  float h3[MAXN][H];
  float c[2 * H];
  float r[H];
  float zr[H];
};

// y for one sample; all activations left in `A` for a backward pass.
// Called by one full wave (blockDim.x == 64).
__device__ float forward_one(const uint16_t* __restrict__ W, const float* __restrict__ x, const uint16_t* adj,
                             uint32_t meta, Acts& A) {
  const int j = threadIdx.x;
  const int k = meta & 0xFF;
  const uint32_t smask = meta >> 16;
  for (int i = j; i < MAXN * FIN; i += H) (&A.h0[0][0])[i] = (i / FIN) < k ? x[i] : 0.f;
  __syncthreads();
  // three message-passing layers
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int IN = l == 0 ? FIN : H;
    float(*hin)[H] = l == 1 ? A.h1 : A.h2;
    float(*ain)[H] = l == 1 ? A.a1 : A.a2;
    float(*hout)[H] = l == 0 ? A.h1 : (l == 1 ? A.h2 : A.h3);
    const int ows = l == 0 ? O_WS0 : (l == 1 ? O_WS1 : O_WS2);
    const int own = l == 0 ? O_WN0 : (l == 1 ? O_WN1 : O_WN2);
    const int ob = l == 0 ? O_B0 : (l == 1 ? O_B1 : O_B2);
    // neighbour sums: lane j owns input column j
    if (j < IN) {
      for (int v = 0; v < k; ++v) {
        float s = 0.f;
        for (uint32_t m = adj[v]; m; m &= m - 1) {
          const int u = __builtin_ctz(m);
          s += l == 0 ? A.h0[u][j] : hin[u][j];
        }
        if (l == 0) A.a0[v][j] = s;
        else ain[v][j] = s;
      }
    }
    __syncthreads();
    float acc[MAXN];
    const float b = bf(W, ob + j);
#pragma unroll
    for (int v = 0; v < MAXN; ++v) acc[v] = b;
    for (int i = 0; i < IN; ++i) {
      const float ws = bf(W, ows + i * H + j), wn = bf(W, own + i * H + j);
#pragma unroll
      for (int v = 0; v < MAXN; ++v) {
        if (v < k) {
          const float hv = l == 0 ? A.h0[v][i] : hin[v][i];
          const float av = l == 0 ? A.a0[v][i] : ain[v][i];
          acc[v] += ws * hv + wn * av;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < MAXN; ++v) hout[v][j] = v < k ? fmaxf(acc[v], 0.f) : 0.f;
    __syncthreads();
  }
  // readout
  float ss = 0.f, sa = 0.f;
  for (int v = 0; v < k; ++v) {
    sa += A.h3[v][j];
    if (smask >> v & 1) ss += A.h3[v][j];
  }
  A.c[j] = ss;
  A.c[H + j] = sa / (float)k;
  __syncthreads();
  float z = bf(W, O_BR + j);
  for (int i = 0; i < 2 * H; ++i) z += bf(W, O_WR + i * H + j) * A.c[i];
  A.zr[j] = z;
  const float rj = fmaxf(z, 0.f);
  A.r[j] = rj;
  float y = rj * bf(W, O_WO + j);
  for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o);
  y += bf(W, O_BO);
  __syncthreads();
  return y;
}

}  // namespace

__global__ __launch_bounds__(64) void k_mpnn_forward(const uint16_t* __restrict__ W, uint32_t n,
                                                     const float* __restrict__ X, const uint16_t* __restrict__ ADJ,
                                                     const uint32_t* __restrict__ META, float* __restrict__ Y) {
  __shared__ Acts A;
  __shared__ uint16_t adj[MAXN];
  for (uint32_t s = blockIdx.x; s < n; s += gridDim.x) {
    if (threadIdx.x < MAXN) adj[threadIdx.x] = ADJ[(uint64_t)s * MAXN + threadIdx.x];
    __syncthreads();
    const float y = forward_one(W, X + (uint64_t)s * MAXN * FIN, adj, META[s], A);
    if (threadIdx.x == 0) Y[s] = y;
    __syncthreads();
  }
}

// Forward + backward over samples s = blockIdx.x, += gridDim.x; the block's
// gradient sum goes to part[blockIdx.x][NPARAM], its squared error to
// loss[blockIdx.x].  Loss = mean over the n samples of (y - t)^2.
__global__ __launch_bounds__(64) void k_mpnn_backward(const uint16_t* __restrict__ W, const uint16_t* __restrict__ WT,
                                                      uint32_t n, const float* __restrict__ X,
                                                      const uint16_t* __restrict__ ADJ,
                                                      const uint32_t* __restrict__ META,
                                                      const float* __restrict__ T, float* __restrict__ part,
                                                      float* __restrict__ loss) {
  float* G = part + (uint64_t)blockIdx.x * NPARAM;  // this block's gradient sum
  __shared__ Acts A;
  __shared__ float dh[MAXN][H], dz[MAXN][H], dag[MAXN][H];
  __shared__ float dzr[H];
  __shared__ uint16_t adj[MAXN];
  const int j = threadIdx.x;
  for (int i = j; i < NPARAM; i += H) G[i] = 0.f;
  float lsum = 0.f;
  const float inv_n = 1.f / (float)n;
  __syncthreads();
  for (uint32_t s = blockIdx.x; s < n; s += gridDim.x) {
    if (j < MAXN) adj[j] = ADJ[(uint64_t)s * MAXN + j];
    __syncthreads();
    const uint32_t meta = META[s];
    const int k = meta & 0xFF;
    const uint32_t smask = meta >> 16;
    const float y = forward_one(W, X + (uint64_t)s * MAXN * FIN, adj, meta, A);
    const float e = y - T[s];
    if (j == 0) lsum += e * e;
    const float dy = 2.f * e * inv_n;
    // output layer
    G[O_WO + j] += dy * A.r[j];
    if (j == 0) G[O_BO] += dy;
    const float dzrj = A.zr[j] > 0.f ? dy * bf(W, O_WO + j) : 0.f;
    dzr[j] = dzrj;
    G[O_BR + j] += dzrj;
    for (int i = 0; i < 2 * H; ++i) G[O_WR + i * H + j] += dzrj * A.c[i];
    __syncthreads();
    // dc[i] = sum_j Wr[i][j] dzr[j]; transposed copy WT holds Wr^T[j][i]
    float dcS = 0.f, dcA = 0.f;
    for (int jj = 0; jj < H; ++jj) {
      dcS += bf(WT, O_WR + jj * 2 * H + j) * dzr[jj];
      dcA += bf(WT, O_WR + jj * 2 * H + H + j) * dzr[jj];
    }
    for (int v = 0; v < MAXN; ++v) dh[v][j] = v < k ? ((smask >> v & 1) ? dcS : 0.f) + dcA / (float)k : 0.f;
    __syncthreads();
    // message-passing layers, last to first
    for (int l = 2; l >= 0; --l) {
      const int IN = l == 0 ? FIN : H;
      float(*hout)[H] = l == 0 ? A.h1 : (l == 1 ? A.h2 : A.h3);
      float(*hin)[H] = l == 1 ? A.h1 : A.h2;
      float(*ain)[H] = l == 1 ? A.a1 : A.a2;
      const int ows = l == 0 ? O_WS0 : (l == 1 ? O_WS1 : O_WS2);
      const int own = l == 0 ? O_WN0 : (l == 1 ? O_WN1 : O_WN2);
      const int ob = l == 0 ? O_B0 : (l == 1 ? O_B1 : O_B2);
      float dbj = 0.f;
      for (int v = 0; v < k; ++v) {
        const float d = hout[v][j] > 0.f ? dh[v][j] : 0.f;
        dz[v][j] = d;
        dbj += d;
      }
      for (int v = k; v < MAXN; ++v) dz[v][j] = 0.f;
      G[ob + j] += dbj;
      __syncthreads();
      // weight gradients: lane j owns column j
      for (int i = 0; i < IN; ++i) {
        float gs = 0.f, gn = 0.f;
        for (int v = 0; v < k; ++v) {
          const float hv = l == 0 ? A.h0[v][i] : hin[v][i];
          const float av = l == 0 ? A.a0[v][i] : ain[v][i];
          gs += dz[v][j] * hv;
          gn += dz[v][j] * av;
        }
        G[ows + i * H + j] += gs;
        G[own + i * H + j] += gn;
      }
      if (l > 0) {
        // input gradients: lane i (= j) owns input column i; W^T[o][i]
        float ds[MAXN], dn[MAXN];
#pragma unroll
        for (int v = 0; v < MAXN; ++v) ds[v] = dn[v] = 0.f;
        for (int o = 0; o < H; ++o) {
          const float wts = bf(WT, ows + o * H + j), wtn = bf(WT, own + o * H + j);
#pragma unroll
          for (int v = 0; v < MAXN; ++v) {
            if (v < k) {
              ds[v] += wts * dz[v][o];
              dn[v] += wtn * dz[v][o];
            }
          }
        }
#pragma unroll
        for (int v = 0; v < MAXN; ++v) dag[v][j] = dn[v];
        __syncthreads();
        for (int v = 0; v < MAXN; ++v) {
          float d = ds[v];
          if (v < k)
            for (uint32_t m = adj[v]; m; m &= m - 1) d += dag[__builtin_ctz(m)][j];
          dh[v][j] = v < k ? d : 0.f;
        }
      }
      __syncthreads();
    }
  }
  if (j == 0) loss[blockIdx.x] = lsum * inv_n;
}

__global__ __launch_bounds__(256) void k_mpnn_reduce(const float* __restrict__ part, uint32_t nb,
                                                     float* __restrict__ grad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint32_t)NPARAM) return;
  float s = 0.f;
  for (uint32_t b = 0; b < nb; ++b) s += part[(uint64_t)b * NPARAM + i];
  grad[i] = s;
}

// Adam step on the FP32 master copy (Kingma & Ba, 2015), then the two BF16
// copies.  t is the 1-based step number.
__global__ __launch_bounds__(256) void k_mpnn_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                   const float* __restrict__ g, float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint32_t)NPARAM) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies: W (input-major, as stored) and WT (each matrix transposed).
__global__ __launch_bounds__(256) void k_mpnn_pack(const float* __restrict__ p, uint16_t* __restrict__ W,
                                                   uint16_t* __restrict__ WT) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint32_t)NPARAM) return;
  const uint16_t b = to_bf16(p[i]);
  W[i] = b;
  // find the matrix containing i and its shape (rows = inputs, cols = outputs)
  int base = -1, rows = 0, cols = H;
  if (i < (uint32_t)O_WN0) { base = O_WS0; rows = FIN; }
  else if (i < (uint32_t)O_B0) { base = O_WN0; rows = FIN; }
  else if (i >= (uint32_t)O_WS1 && i < (uint32_t)O_WN1) { base = O_WS1; rows = H; }
  else if (i >= (uint32_t)O_WN1 && i < (uint32_t)O_B1) { base = O_WN1; rows = H; }
  else if (i >= (uint32_t)O_WS2 && i < (uint32_t)O_WN2) { base = O_WS2; rows = H; }
  else if (i >= (uint32_t)O_WN2 && i < (uint32_t)O_B2) { base = O_WN2; rows = H; }
  else if (i >= (uint32_t)O_WR && i < (uint32_t)O_BR) { base = O_WR; rows = 2 * H; }
  if (base < 0) {
    WT[i] = b;
    return;
  }
  const uint32_t off = i - base, r = off / cols, c = off % cols;  // W[r][c]
  WT[base + c * rows + r] = b;
}

// ------------------------------------------------------------------ host
namespace {
template <typename T>
T* malloc_dev(uint64_t n) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(-5, "mpnn: device allocation failed");
  }
  return (T*)p;
}
}  // namespace

Mpnn::~Mpnn() {
  if (ctx) (void)hipSetDevice(ctx->device);
  for (void* q : {(void*)p, (void*)m, (void*)v, (void*)g, (void*)w, (void*)wt, (void*)part, (void*)loss,
                  (void*)x, (void*)adj, (void*)meta, (void*)t, (void*)y})
    if (q) (void)hipFree(q);
}

void mpnn_init(Context& c, Mpnn& n, const float* params) {
  HIPCHECK(hipSetDevice(c.device));
  n.ctx = &c;
  n.p = malloc_dev<float>(NPARAM);
  n.m = malloc_dev<float>(NPARAM);
  n.v = malloc_dev<float>(NPARAM);
  n.g = malloc_dev<float>(NPARAM);
  n.w = malloc_dev<uint16_t>(NPARAM);
  n.wt = malloc_dev<uint16_t>(NPARAM);
  n.nblocks = (uint32_t)c.cus;
  n.part = malloc_dev<float>((uint64_t)n.nblocks * NPARAM);
  n.loss = malloc_dev<float>(n.nblocks);
  HIPCHECK(hipMemset(n.m, 0, NPARAM * 4));
  HIPCHECK(hipMemset(n.v, 0, NPARAM * 4));
  mpnn_set(n, params);
}

void mpnn_set(Mpnn& n, const float* params) {
  HIPCHECK(hipSetDevice(n.ctx->device));
  HIPCHECK(hipMemcpy(n.p, params, NPARAM * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mpnn_pack, dim3((NPARAM + 255) / 256), dim3(256), 0, n.ctx->stream, n.p, n.w, n.wt);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(n.ctx->stream));
  n.step = 0;
  HIPCHECK(hipMemset(n.m, 0, NPARAM * 4));
  HIPCHECK(hipMemset(n.v, 0, NPARAM * 4));
}

void mpnn_get(Mpnn& n, float* params) {
  HIPCHECK(hipSetDevice(n.ctx->device));
  HIPCHECK(hipMemcpy(params, n.p, NPARAM * 4, hipMemcpyDeviceToHost));
}

static void upload(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta, const float* t) {
  if (cnt > n.cap) {
    for (void* q : {(void*)n.x, (void*)n.adj, (void*)n.meta, (void*)n.t, (void*)n.y})
      if (q) (void)hipFree(q);
    n.cap = std::max<uint32_t>(cnt, 1024);
    n.x = malloc_dev<float>((uint64_t)n.cap * MAXN * FIN);
    n.adj = malloc_dev<uint16_t>((uint64_t)n.cap * MAXN);
    n.meta = malloc_dev<uint32_t>(n.cap);
    n.t = malloc_dev<float>(n.cap);
    n.y = malloc_dev<float>(n.cap);
  }
  hipStream_t s = n.ctx->stream;
  HIPCHECK(hipMemcpyAsync(n.x, x, (uint64_t)cnt * MAXN * FIN * 4, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(n.adj, adj, (uint64_t)cnt * MAXN * 2, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(n.meta, meta, (uint64_t)cnt * 4, hipMemcpyHostToDevice, s));
  if (t) HIPCHECK(hipMemcpyAsync(n.t, t, (uint64_t)cnt * 4, hipMemcpyHostToDevice, s));
}

static void check_meta(uint32_t cnt, const uint32_t* meta) {
  for (uint32_t i = 0; i < cnt; ++i) {
    const uint32_t k = meta[i] & 0xFF;
    if (k < 1 || k > (uint32_t)MAXN) throw Error(-1, "mpnn: sample with k outside 1..16");
  }
}

void mpnn_forward(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta, float* y) {
  HIPCHECK(hipSetDevice(n.ctx->device));
  if (!cnt) return;
  check_meta(cnt, meta);
  upload(n, cnt, x, adj, meta, nullptr);
  const uint32_t grid = std::min<uint32_t>(cnt, (uint32_t)n.ctx->cus * 16);
  hipLaunchKernelGGL(k_mpnn_forward, dim3(grid), dim3(64), 0, n.ctx->stream, n.w, cnt, n.x, n.adj, n.meta, n.y);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(y, n.y, (uint64_t)cnt * 4, hipMemcpyDeviceToHost, n.ctx->stream));
  HIPCHECK(hipStreamSynchronize(n.ctx->stream));
}

double mpnn_gradient(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta,
                     const float* t, float* grad_out) {
  HIPCHECK(hipSetDevice(n.ctx->device));
  if (!cnt) throw Error(-1, "mpnn: empty batch");
  check_meta(cnt, meta);
  upload(n, cnt, x, adj, meta, t);
  const uint32_t nb = std::min<uint32_t>(n.nblocks, cnt);
  hipLaunchKernelGGL(k_mpnn_backward, dim3(nb), dim3(64), 0, n.ctx->stream, n.w, n.wt, cnt, n.x, n.adj, n.meta, n.t,
                     n.part, n.loss);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mpnn_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, n.ctx->stream, n.part, nb, n.g);
  HIPCHECK(hipGetLastError());
  std::vector<float> l(nb);
  HIPCHECK(hipMemcpyAsync(l.data(), n.loss, nb * 4, hipMemcpyDeviceToHost, n.ctx->stream));
  if (grad_out) HIPCHECK(hipMemcpyAsync(grad_out, n.g, NPARAM * 4, hipMemcpyDeviceToHost, n.ctx->stream));
  HIPCHECK(hipStreamSynchronize(n.ctx->stream));
  double s = 0;
  for (float v : l) s += v;
  return s;
}

double mpnn_train_step(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta,
                       const float* t, float lr) {
  const double loss = mpnn_gradient(n, cnt, x, adj, meta, t, nullptr);
  n.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)n.step), c2 = 1.f - std::pow(b2, (float)n.step);
  hipLaunchKernelGGL(k_mpnn_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, n.ctx->stream, n.p, n.m, n.v, n.g, lr, b1,
                     b2, 1e-8f, c1, c2);
  hipLaunchKernelGGL(k_mpnn_pack, dim3((NPARAM + 255) / 256), dim3(256), 0, n.ctx->stream, n.p, n.w, n.wt);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(n.ctx->stream));
  return loss;
}

}  // namespace cuts

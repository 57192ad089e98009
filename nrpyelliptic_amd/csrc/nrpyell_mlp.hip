// nrpyell_mlp.hip -- the learned initial guess: a coordinate MLP u(x; punctures), trained on
// solutions this solver produced, run on hand-written BF16 MFMA kernels (CDNA4
// v_mfma_f32_32x32x16_bf16) with FP32 accumulation. No reference counterpart.
//
// Network: K0 = 32 (padded) input features -> 5 x [Linear(256) + SiLU] -> Linear(1).
// Activations between layers are stored in BF16 (MFMA operands), pre-activations in FP32;
// master weights, gradients and Adam moments are FP32, weights are re-rounded to BF16 after
// every update.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <string.h>

#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "nrpyell_internal.h"

typedef __bf16 bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

#define MLP_K0 64  // input features, zero-padded so every GEMM dimension is a multiple of 64
#define MLP_NFEAT 23
#define MLP_W 256
#define MLP_HIDDEN 5  // hidden layers of width 256 (6 linear layers in total)

// Parameter offsets (elements) of W_l and b_l, l = 1..5, and of the output layer (index 6).
struct MlpOffsets {
  long w[MLP_HIDDEN + 2], b[MLP_HIDDEN + 2];
};

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ bf16_t to_bf16(float x) { return (bf16_t)x; }

// C(m,c) = sum_i A(m,i) B(c,i) with A(m,i) = A[m*a_m + i*a_i], B(c,i) = B[c*b_c + i*b_i].
// Block tile 64 x 64 (4 waves, each 32 x 32 = one MFMA accumulator), inner dimension in
// chunks of 32 staged through LDS. Requirements (checked by the host): M, N multiples of 64,
// the inner range [i_begin, i_end) a multiple of 32.
//   EPI_FWD : z = acc + bias[c]; Z[m*N+c] = z; H[m*N+c] = bf16(silu(z))
//   EPI_BWD : out[m*N+c] = bf16(acc * silu'(Zprev[m*N+c]))   (gradient w.r.t. the previous pre-activation)
//   EPI_WGT : atomicAdd(G[m*N+c], acc)                       (weight gradient, inner dim split over blockIdx.z)
enum { EPI_FWD = 0, EPI_BWD = 1, EPI_WGT = 2 };

template <int EPI>
__global__ void __launch_bounds__(256) mfma_gemm_kernel(const bf16_t *__restrict__ A, long a_m, long a_i,
                                                        const bf16_t *__restrict__ B, long b_c, long b_i, int M,
                                                        int N, int inner, int chunk, const float *__restrict__ bias,
                                                        float *__restrict__ Z, bf16_t *__restrict__ H,
                                                        const float *__restrict__ Zprev, float *__restrict__ G) {
  __shared__ bf16_t As[64][32 + 8];
  __shared__ bf16_t Bs[64][32 + 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int i_begin = blockIdx.z * chunk;
  int i_end = i_begin + chunk;
  if (i_end > inner) i_end = inner;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  const int wm = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 31, lj = lane >> 5;
  for (int ib = i_begin; ib < i_end; ib += 32) {
    // stage 64 x 32 of A and of B: 2048 elements each, 8 per thread
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = tid + 256 * q;
      const int row = e >> 5, col = e & 31;
      As[row][col] = A[(long)(m0 + row) * a_m + (long)(ib + col) * a_i];
      Bs[row][col] = B[(long)(c0 + row) * b_c + (long)(ib + col) * b_i];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; kk += 16) {
      bf16x8 a, b;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        a[t] = As[wm + li][kk + 8 * lj + t];
        b[t] = Bs[wc + li][kk + 8 * lj + t];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator r of lane (li, lj): row 8*(r/4) + 4*lj + r%4, column li
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int m = m0 + wm + 8 * (r >> 2) + 4 * lj + (r & 3);
    const int c = c0 + wc + li;
    const long o = (long)m * N + c;
    if (EPI == EPI_FWD) {
      const float z = acc[r] + bias[c];
      Z[o] = z;
      H[o] = to_bf16(z * sigmoidf_(z));
    } else if (EPI == EPI_BWD) {
      const float zp = Zprev[o];
      const float s = sigmoidf_(zp);
      H[o] = to_bf16(acc[r] * (s * (1.0f + zp * (1.0f - s))));
    } else {
      atomicAdd(&G[o], acc[r]);
    }
  }
}

// Inference, all six layers in one launch: a block takes 64 rows through the five hidden
// layers with the activations held in LDS (two 64 x 256 BF16 buffers, ping-pong), so nothing
// but the 64 input features and the output leave the chip; the pre-activations are not kept
// (training uses the per-layer kernels, whose backward needs them). Each of the 4 waves owns 64
// output columns: 2 x 2 MFMA accumulators over the 64 rows; the weights (544 KB of BF16, shared
// by every block) are read from L2 as 16-byte MFMA operands. Same arithmetic as the per-layer
// path: FP32 accumulation over BF16 operands in ascending K, z = acc + b, h = bf16(silu(z)),
// output y = sum_k h_k w_k + b in FP32.
constexpr int FI_R = 64, FI_LD = MLP_W + 8;  // rows per block; LDS row stride (528 B)

__global__ void __launch_bounds__(256) mlp_forward_fused_kernel(const bf16_t *__restrict__ X, const bf16_t *__restrict__ pw,
                                                                const float *__restrict__ p, MlpOffsets off, int M,
                                                                float *__restrict__ y) {
  __shared__ __attribute__((aligned(16))) bf16_t Hbuf[2][FI_R * FI_LD];
  __shared__ float s_out[4][FI_R];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 31, lj = lane >> 5;
  const int r0 = blockIdx.x * FI_R;
  // input features: 64 rows x 64 BF16 = 512 chunks of 16 B, 2 per thread (rows past M clamped)
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = tid + 256 * q, row = e >> 3, ch = e & 7;
    const int m = min(r0 + row, M - 1);
    *(bf16x8 *)(&Hbuf[0][row * FI_LD + 8 * ch]) = *(const bf16x8 *)(X + (long)m * MLP_K0 + 8 * ch);
  }
  __syncthreads();
  int cur = 0;
  const int wc = wave * 64;
  for (int l = 1; l <= MLP_HIDDEN; l++) {
    const int Kl = l == 1 ? MLP_K0 : MLP_W;
    const bf16_t *W = pw + off.w[l];  // W_l [256 x K_l]
    floatx16 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[x][yy][r] = 0.0f;
    const bf16_t *Hin = Hbuf[cur];
    for (int k = 0; k < Kl; k += 16) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int x = 0; x < 2; x++) a[x] = *(const bf16x8 *)(Hin + (32 * x + li) * FI_LD + k + 8 * lj);
#pragma unroll
      for (int yy = 0; yy < 2; yy++) b[yy] = *(const bf16x8 *)(W + (long)(wc + 32 * yy + li) * Kl + k + 8 * lj);
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int yy = 0; yy < 2; yy++)
          acc[x][yy] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x], b[yy], acc[x][yy], 0, 0, 0);
    }
    bf16_t *Hout = Hbuf[cur ^ 1];
    const float *bias = p + off.b[l];
    // accumulator r of lane (li, lj) in tile (x, yy): row 8*(r/4) + 4*lj + r%4, column li
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int yy = 0; yy < 2; yy++) {
        const int c = wc + 32 * yy + li;
        const float bc = bias[c];
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int row = 32 * x + 8 * (r >> 2) + 4 * lj + (r & 3);
          const float z = acc[x][yy][r] + bc;
          Hout[row * FI_LD + c] = to_bf16(z * sigmoidf_(z));
        }
      }
    __syncthreads();
    cur ^= 1;
  }
  // output layer: wave w sums columns 64w..64w+63 of every row, then the 4 partials are added
  const bf16_t *Hl = Hbuf[cur];
  const float *wout = p + off.w[MLP_HIDDEN + 1];
  for (int row = lane; row < FI_R; row += 64) {
    float acc = 0.0f;
#pragma unroll 8
    for (int c = 0; c < 64; c++) acc += (float)Hl[row * FI_LD + wc + c] * wout[wc + c];
    s_out[wave][row] = acc;
  }
  __syncthreads();
  if (tid < FI_R && r0 + tid < M)
    y[r0 + tid] = ((s_out[0][tid] + s_out[1][tid]) + (s_out[2][tid] + s_out[3][tid])) + p[off.b[MLP_HIDDEN + 1]];
}

// Output layer (256 -> 1): y[m] = sum_k H[m][k] w[k] + b (FP32 dot over BF16 activations).
__global__ void out_layer_kernel(const bf16_t *__restrict__ H, const float *__restrict__ w, const float *__restrict__ b,
                                 int M, float *__restrict__ y) {
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (m >= M) return;
  float acc = 0.0f;
#pragma unroll
  for (int k = lane; k < MLP_W; k += 64) acc += (float)H[(long)m * MLP_W + k] * w[k];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane == 0) y[m] = acc + b[0];
}

// Loss and the output-layer backward: dy = 2 (y - t)/M per row, dZ5 = dy w silu'(Z5); the
// loss is reduced per block (one atomic per block).
__global__ void out_backward_kernel(const float *__restrict__ Z5, const float *__restrict__ w,
                                    const float *__restrict__ y, const float *__restrict__ t, int M,
                                    float *__restrict__ dy_out, bf16_t *__restrict__ dZ, double *__restrict__ loss) {
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  __shared__ double sl[4];
  double l = 0.0;
  if (m < M) {
    const float diff = y[m] - t[m];
    const float dy = 2.0f * diff / (float)M;
    for (int k = lane; k < MLP_W; k += 64) {
      const long o = (long)m * MLP_W + k;
      const float z = Z5[o], s = sigmoidf_(z);
      dZ[o] = to_bf16(dy * w[k] * (s * (1.0f + z * (1.0f - s))));
    }
    if (lane == 0) {
      dy_out[m] = dy;
      l = (double)diff * (double)diff / (double)M;
    }
  }
  if (lane == 0) sl[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (sl[0] + sl[1]) + (sl[2] + sl[3]));
}

// Output-layer weight gradients: gw[k] = sum_m dy[m] H[m][k], gb = sum_m dy[m]; rows split
// over blockIdx.y, one atomic per column per block.
__global__ void out_wgrad_kernel(const bf16_t *__restrict__ H, const float *__restrict__ dy, int M,
                                 float *__restrict__ gw, float *__restrict__ gb) {
  const int k = threadIdx.x;  // 256 columns
  const int rows = (M + gridDim.y - 1) / gridDim.y, r0 = blockIdx.y * rows;
  float acc = 0.0f, accb = 0.0f;
  for (int m = r0; m < r0 + rows && m < M; m++) {
    const float d = dy[m];
    acc += d * (float)H[(long)m * MLP_W + k];
    accb += d;
  }
  atomicAdd(&gw[k], acc);
  if (k == 0) atomicAdd(gb, accb);
}

// Column sums of a BF16 [M x N] matrix into an FP32 gradient (bias gradients).
__global__ void colsum_kernel(const bf16_t *__restrict__ X, int M, int N, float *__restrict__ g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = (M + gridDim.y - 1) / gridDim.y, r0 = blockIdx.y * rows;
  if (c >= N) return;
  float acc = 0.0f;
  for (int m = r0; m < r0 + rows && m < M; m++) acc += (float)X[(long)m * N + c];
  atomicAdd(&g[c], acc);
}

__global__ void adam_kernel(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ mo,
                            float *__restrict__ ve, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float m = b1 * mo[k] + (1.0f - b1) * gk;
  const float v = b2 * ve[k] + (1.0f - b2) * gk * gk;
  mo[k] = m;
  ve[k] = v;
  p[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void to_bf16_kernel(long n, const float *__restrict__ x, bf16_t *__restrict__ y) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = to_bf16(x[k]);
}

__global__ void gather_rows_kernel(const bf16_t *__restrict__ X, const float *__restrict__ t, long N, int M,
                                   unsigned long long seed, bf16_t *__restrict__ Xb, float *__restrict__ tb) {
  const int m = blockIdx.x;
  // splitmix64 of (seed, m): uniform row index
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(m + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const long row = (long)(z % (unsigned long long)N);
  for (int k = threadIdx.x; k < MLP_K0; k += blockDim.x) Xb[(long)m * MLP_K0 + k] = X[row * MLP_K0 + k];
  if (threadIdx.x == 0) tb[m] = t[row];
}

__global__ void zero_kernel(long n, float *x) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = 0.0f;
}

// ---------------------------------------------------------------------------------------
// Features of the learned initial guess at every interior and pure outer point of a grid.
// With b the SinhSymTP focus (bScale), r = |X|, r_n = |X - x_n|:
//   point: X/(r+b), Y/(r+b), Z/(r+b), r/(r+b), log(1+r/b)/10, b/(r_0+b), b/(r_1+b)
//   configuration: m_0, m_1, D/20, 10 P_0, 10 P_1, S_0/m_0^2, S_1/m_1^2, b/10
// (23 numbers, zero-padded to 64). The network predicts t = log(50 u (r+b)): u > 0 everywhere
// (lap u = -AA^2/(8 psi^7) <= 0 with u -> 0 far away), and u (r+b) stays within a few
// decades over the grid, so t is O(1) at every point; u = exp(t)/(50 (r+b)).
// ---------------------------------------------------------------------------------------
struct FeatureParams {
  double AMAX, w, b, D;
  double m[2], pos[2][3], P[2][3], S[2][3];
};

__device__ inline void point_features(const FeatureParams &f, double x0, double x1, double x2, float out[MLP_K0],
                                      double &scale) {
  const double f0 = f.AMAX * (exp(x0 / f.w) - exp(-x0 / f.w)) / f.D;
  const double X = f0 * sin(x1) * cos(x2), Y = f0 * sin(x1) * sin(x2), Z = sqrt(f0 * f0 + f.b * f.b) * cos(x1);
  const double r = sqrt(X * X + Y * Y + Z * Z), rb = r + f.b;
  double rn[2];
  for (int n = 0; n < 2; n++) {
    const double dx = X - f.pos[n][0], dy = Y - f.pos[n][1], dz = Z - f.pos[n][2];
    rn[n] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  const double sep = sqrt((f.pos[0][0] - f.pos[1][0]) * (f.pos[0][0] - f.pos[1][0]) +
                          (f.pos[0][1] - f.pos[1][1]) * (f.pos[0][1] - f.pos[1][1]) +
                          (f.pos[0][2] - f.pos[1][2]) * (f.pos[0][2] - f.pos[1][2]));
  int k = 0;
  out[k++] = (float)(X / rb);
  out[k++] = (float)(Y / rb);
  out[k++] = (float)(Z / rb);
  out[k++] = (float)(r / rb);
  out[k++] = (float)(log1p(r / f.b) / 10.0);
  out[k++] = (float)(f.b / (rn[0] + f.b));
  out[k++] = (float)(f.b / (rn[1] + f.b));
  out[k++] = (float)f.m[0];
  out[k++] = (float)f.m[1];
  out[k++] = (float)(sep / 20.0);
  for (int n = 0; n < 2; n++)
    for (int d = 0; d < 3; d++) out[k++] = (float)(10.0 * f.P[n][d]);
  for (int n = 0; n < 2; n++)
    for (int d = 0; d < 3; d++) out[k++] = (float)(f.S[n][d] / fmax(f.m[n] * f.m[n], 1e-12));
  out[k++] = (float)(f.b / 10.0);
  while (k < MLP_K0) out[k++] = 0.0f;
  scale = 1.0 / (50.0 * rb);
}

// One thread per interior or pure-outer point (i0 in [NG, G0), i1, i2 interior).
__global__ void features_kernel(GridDims g, FeatureParams f, const double *xx0, const double *xx1, const double *xx2,
                                bf16_t *__restrict__ X, double *__restrict__ scale, long *__restrict__ idx) {
  const int n0 = g.G0 - NRPYELL_NG;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npts = (long)n0 * g.N1 * g.N2;
  if (k >= npts) return;
  const int i0 = NRPYELL_NG + (int)(k % n0);
  const int i1 = NRPYELL_NG + (int)((k / n0) % g.N1);
  const int i2 = NRPYELL_NG + (int)(k / ((long)n0 * g.N1));
  float out[MLP_K0];
  double s;
  point_features(f, xx0[i0], xx1[i1], xx2[i2], out, s);
#pragma unroll
  for (int q = 0; q < MLP_K0; q++) X[k * MLP_K0 + q] = to_bf16(out[q]);
  scale[k] = s;
  idx[k] = idx3(g, i0, i1, i2);
}

// u = t * scale at the listed points, v = eta u, and the parity images of each point.
__global__ void seed_state_kernel(GridDims g, long npts, const float *__restrict__ t, const double *__restrict__ scale,
                                  const long *__restrict__ idx, double eta, double *__restrict__ yn) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const double u = exp((double)t[k]) * scale[k];
  const long p = idx[k];
  yn[p] = u;
  yn[g.ntot + p] = eta * u;
  const int i0 = (int)(p % g.G0), i1 = (int)((p / g.G0) % g.G1), i2 = (int)(p / ((long)g.G0 * g.G1));
  for_each_image(g, i0, i1, i2, [&](long d) {
    yn[d] = u;
    yn[g.ntot + d] = eta * u;
  });
}

// One binomial smoothing sub-pass of u along direction DIR over the interior points,
// out = (u[-1] + 2 u + u[+1]) / 4 (phi neighbours through the periodic map), pure outer points
// copied; every point writes its interior-plane parity images, so the next sub-pass reads
// current ghosts. It removes the grid-scale part of the network's pointwise error (BF16
// activations round independently at neighbouring points), which the Laplacian in the
// residual would otherwise amplify by 1/dx^2.
template <int DIR>
__global__ void smooth_kernel(GridDims g, const double *__restrict__ in, double *__restrict__ out) {
  const int i0 = NRPYELL_NG + blockIdx.x * blockDim.x + threadIdx.x;
  const int i1 = NRPYELL_NG + blockIdx.y, i2 = NRPYELL_NG + blockIdx.z;
  if (i0 >= g.G0) return;
  const long p = idx3(g, i0, i1, i2);
  double v = in[p];
  if (i0 < g.N0 + NRPYELL_NG) {
    long pm, pp;
    if (DIR == 0) { pm = p - 1; pp = p + 1; }
    else if (DIR == 1) { pm = p - g.G0; pp = p + g.G0; }
    else { pm = idx3(g, i0, i1, wrap2(g, i2 - 1)); pp = idx3(g, i0, i1, wrap2(g, i2 + 1)); }
    v = 0.25 * (in[pm] + in[pp]) + 0.5 * v;
  }
  out[p] = v;
  for_each_image_in_planes(g, i0, i1, i2, [&](long d) { out[d] = v; });
}

__global__ void v_from_u_kernel(long n, double eta, double *y) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[n + k] = eta * y[k];
}

// log(u / scale) at the listed points: the network's training target from a solved state
// (u / scale floored at 1e-8).
__global__ void target_kernel(long npts, const double *__restrict__ yn, const double *__restrict__ scale,
                              const long *__restrict__ idx, float *__restrict__ t) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < npts) t[k] = (float)log(fmax(yn[idx[k]] / scale[k], 1e-8));
}

// ---------------------------------------------------------------------------------------
// Host object
// ---------------------------------------------------------------------------------------
template <typename T>
static T *mlp_alloc(size_t n) {
  void *p = nullptr;
  HIPCHK(hipMalloc(&p, n * sizeof(T) > 0 ? n * sizeof(T) : 1));
  return (T *)p;
}

struct nrpyell_mlp {
  // parameter layout (FP32, concatenated): for l = 1..5: W_l [256 x K_l], b_l [256]; then w_out [256], b_out [1]
  std::vector<long> w_off, b_off, w_size;
  long n_params = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  bf16_t *pw = nullptr;  // BF16 copy of all parameters (weights read by the MFMA kernels)
  long adam_t = 0;
  int cap = 0;           // activations allocated for this many rows
  bf16_t *Hs[MLP_HIDDEN + 1] = {};  // H_0 = input, H_1..H_5
  float *Zs[MLP_HIDDEN + 1] = {};   // Z_1..Z_5
  bf16_t *dZa = nullptr, *dZb = nullptr;
  float *y = nullptr, *tb = nullptr;
  double *loss = nullptr;
  hipStream_t st = nullptr;
  ~nrpyell_mlp() {
    (void)hipFree(p); (void)hipFree(g); (void)hipFree(m); (void)hipFree(v); (void)hipFree(pw);
    for (auto *h : Hs) (void)hipFree(h);
    for (auto *z : Zs) (void)hipFree(z);
    (void)hipFree(dZa); (void)hipFree(dZb); (void)hipFree(y); (void)hipFree(tb); (void)hipFree(loss); (void)hipFree(dy);
    if (st) (void)hipStreamDestroy(st);
  }
  int K(int l) const { return l == 1 ? MLP_K0 : MLP_W; }
  void ensure(int M) {
    if (M <= cap) return;
    // release first; a failed allocation below (HipFailure) leaves cap 0 and null or valid
    // pointers only, so the next call allocates again
    cap = 0;
    for (auto *&h : Hs) { (void)hipFree(h); h = nullptr; }
    for (auto *&z : Zs) { (void)hipFree(z); z = nullptr; }
    (void)hipFree(dZa); (void)hipFree(dZb); (void)hipFree(y); (void)hipFree(tb); (void)hipFree(dy);
    dZa = dZb = nullptr;
    y = tb = dy = nullptr;
    Hs[0] = mlp_alloc<bf16_t>((size_t)M * MLP_K0);
    for (int l = 1; l <= MLP_HIDDEN; l++) {
      Hs[l] = mlp_alloc<bf16_t>((size_t)M * MLP_W);
      Zs[l] = mlp_alloc<float>((size_t)M * MLP_W);
    }
    dZa = mlp_alloc<bf16_t>((size_t)M * MLP_W);
    dZb = mlp_alloc<bf16_t>((size_t)M * MLP_W);
    y = mlp_alloc<float>(M);
    tb = mlp_alloc<float>(M);
    dy = mlp_alloc<float>(M);
    cap = M;
  }
  void refresh_bf16() {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((n_params + 255) / 256), dim3(256), 0, st, n_params, p, pw);
  }
  // forward from Hs[0] (M rows, M a multiple of 64) to y
  void forward(int M) {
    for (int l = 1; l <= MLP_HIDDEN; l++) {
      const int Kl = K(l);
      dim3 grd(M / 64, MLP_W / 64, 1);
      hipLaunchKernelGGL(mfma_gemm_kernel<EPI_FWD>, grd, dim3(256), 0, st, Hs[l - 1], (long)Kl, 1L,
                         pw + w_off[l], (long)Kl, 1L, M, MLP_W, Kl, Kl, p + b_off[l], Zs[l], Hs[l], nullptr,
                         nullptr);
    }
    hipLaunchKernelGGL(out_layer_kernel, dim3((M + 3) / 4), dim3(256), 0, st, Hs[MLP_HIDDEN], p + w_off[6],
                       p + b_off[6], M, y);
  }
  // loss and gradients into g (zeroed first); targets in tb
  void backward(int M) {
    hipLaunchKernelGGL(zero_kernel, dim3((n_params + 255) / 256), dim3(256), 0, st, n_params, g);
    HIPCHK(hipMemsetAsync(loss, 0, sizeof(double), st));
    hipLaunchKernelGGL(out_backward_kernel, dim3((M + 3) / 4), dim3(256), 0, st, Zs[MLP_HIDDEN], p + w_off[6], y, tb,
                       M, dy, dZa, loss);
    hipLaunchKernelGGL(out_wgrad_kernel, dim3(1, 64), dim3(MLP_W), 0, st, Hs[MLP_HIDDEN], dy, M, g + w_off[6],
                       g + b_off[6]);
    bf16_t *dZ = dZa, *dZn = dZb;
    for (int l = MLP_HIDDEN; l >= 1; l--) {
      const int Kl = K(l);
      // dW_l[n][k] = sum_m dZ[m][n] H_{l-1}[m][k]: rows n, columns k, inner m split in chunks
      const int chunk = 2048;
      dim3 gw(MLP_W / 64, Kl / 64, (M + chunk - 1) / chunk);
      hipLaunchKernelGGL(mfma_gemm_kernel<EPI_WGT>, gw, dim3(256), 0, st, dZ, 1L, (long)MLP_W, Hs[l - 1], 1L,
                         (long)Kl, MLP_W, Kl, M, chunk, nullptr, nullptr, nullptr, nullptr, g + w_off[l]);
      hipLaunchKernelGGL(colsum_kernel, dim3(MLP_W / 256, 64), dim3(256), 0, st, dZ, M, MLP_W, g + b_off[l]);
      if (l > 1) {
        // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1})
        hipLaunchKernelGGL(mfma_gemm_kernel<EPI_BWD>, dim3(M / 64, MLP_W / 64, 1), dim3(256), 0, st, dZ, (long)MLP_W,
                           1L, pw + w_off[l], 1L, (long)Kl, M, MLP_W, MLP_W, MLP_W, nullptr, nullptr, dZn,
                           Zs[l - 1], nullptr);
        bf16_t *tmp = dZ;
        dZ = dZn;
        dZn = tmp;
      }
    }
  }
  float *dy = nullptr;  // per-row output gradient
};

static void backward_full(nrpyell_mlp *s, int M) { s->backward(M); }

extern "C" int nrpyell_mlp_create(unsigned long long seed, nrpyell_mlp **out) {
  return api_call([&]() -> int {
    auto s = std::make_unique<nrpyell_mlp>();
    s->w_off.assign(7, 0);
    s->b_off.assign(7, 0);
    s->w_size.assign(7, 0);
    long off = 0;
    for (int l = 1; l <= MLP_HIDDEN; l++) {
      s->w_off[l] = off;
      s->w_size[l] = (long)MLP_W * s->K(l);
      off += s->w_size[l];
      s->b_off[l] = off;
      off += MLP_W;
    }
    s->w_off[6] = off;
    s->w_size[6] = MLP_W;
    off += MLP_W;
    s->b_off[6] = off;
    off += 1;
    s->n_params = off;
    // He-style initialisation (uniform, splitmix64 stream), biases zero
    std::vector<float> h(off, 0.0f);
    unsigned long long z = seed;
    auto rnd = [&]() {
      z += 0x9E3779B97F4A7C15ull;
      unsigned long long x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      return (double)(x >> 11) / 9007199254740992.0;
    };
    for (int l = 1; l <= 6; l++) {
      const int fan_in = l == 1 ? MLP_NFEAT : MLP_W;
      const double a = sqrt(6.0 / fan_in) * (l == 6 ? 0.1 : 1.0);
      for (long k = 0; k < s->w_size[l]; k++) h[s->w_off[l] + k] = (float)((2 * rnd() - 1) * a);
    }
    s->p = mlp_alloc<float>(off);
    s->g = mlp_alloc<float>(off);
    s->m = mlp_alloc<float>(off);
    s->v = mlp_alloc<float>(off);
    s->pw = mlp_alloc<bf16_t>(off);
    s->loss = mlp_alloc<double>(1);
    HIPCHK(hipMemcpy(s->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->m, 0, off * sizeof(float)));
    HIPCHK(hipMemset(s->v, 0, off * sizeof(float)));
    HIPCHK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    s->refresh_bf16();
    HIPCHK(hipStreamSynchronize(s->st));
    *out = s.release();
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_mlp_destroy(nrpyell_mlp *s) {
  return api_call([&]() -> int {
    if (s) {
      (void)hipStreamSynchronize(s->st);
      delete s;
    }
    return NRPYELL_OK;
  });
}

extern "C" long nrpyell_mlp_num_params(nrpyell_mlp *s) { return s->n_params; }

extern "C" int nrpyell_mlp_get_params(nrpyell_mlp *s, float *host) {
  return api_call([&]() -> int {
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(host, s->p, s->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_mlp_set_params(nrpyell_mlp *s, const float *host) {
  return api_call([&]() -> int {
    HIPCHK(hipMemcpy(s->p, host, s->n_params * sizeof(float), hipMemcpyHostToDevice));
    s->refresh_bf16();
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

static int check_rows(int M) {
  if (M <= 0 || M % 64) {
    set_last_error("MLP row count must be a positive multiple of 64");
    return NRPYELL_EINVAL;
  }
  return NRPYELL_OK;
}

// Forward pass on M rows of device BF16 features [M x 32]; y (device, FP32) gets M outputs.
extern "C" int nrpyell_mlp_forward(nrpyell_mlp *s, const void *x, int M, float *y_out) {
  return api_call([&]() -> int {
    if (check_rows(M)) return NRPYELL_EINVAL;
    MlpOffsets off;
    for (int l = 1; l <= MLP_HIDDEN + 1; l++) {
      off.w[l] = s->w_off[l];
      off.b[l] = s->b_off[l];
    }
    hipLaunchKernelGGL(mlp_forward_fused_kernel, dim3((M + FI_R - 1) / FI_R), dim3(256), 0, s->st, (const bf16_t *)x,
                       s->pw, s->p, off, M, y_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

// Loss (mean squared error) and its gradient for M rows; grads copied to host when non-null.
extern "C" int nrpyell_mlp_loss_grad(nrpyell_mlp *s, const void *x, const float *t, int M, double *loss,
                                     float *grads_host) {
  return api_call([&]() -> int {
    if (check_rows(M)) return NRPYELL_EINVAL;
    s->ensure(M);
    HIPCHK(hipMemcpyAsync(s->Hs[0], x, (size_t)M * MLP_K0 * sizeof(bf16_t), hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->tb, t, (size_t)M * sizeof(float), hipMemcpyDeviceToDevice, s->st));
    s->forward(M);
    backward_full(s, M);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(loss, s->loss, sizeof(double), hipMemcpyDeviceToHost, s->st));
    if (grads_host) HIPCHK(hipMemcpyAsync(grads_host, s->g, s->n_params * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

static void adam(nrpyell_mlp *s, float lr) {
  s->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)s->adam_t), c2 = 1.0f - powf(b2, (float)s->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3((s->n_params + 255) / 256), dim3(256), 0, s->st, s->n_params, s->p, s->g, s->m,
                     s->v, lr, b1, b2, 1e-8f, c1, c2);
  s->refresh_bf16();
}

// nsteps of Adam on minibatches of `batch` rows drawn uniformly from a device dataset
// (features BF16 [N x 32], targets FP32 [N]); cosine learning-rate decay from lr to lr/100.
// losses (host, nsteps entries, may be null) receives each step's minibatch loss.
extern "C" int nrpyell_mlp_train(nrpyell_mlp *s, const void *X, const float *t, long N, int batch, int nsteps,
                                 float lr, unsigned long long seed, double *losses) {
  return api_call([&]() -> int {
    if (check_rows(batch) || N <= 0) return NRPYELL_EINVAL;
    s->ensure(batch);
    std::vector<double> hl;
    DevBuf<double> dlb(nsteps > 0 ? nsteps : 1);
    double *dl = dlb.p;
    for (int k = 0; k < nsteps; k++) {
      hipLaunchKernelGGL(gather_rows_kernel, dim3(batch), dim3(32), 0, s->st, (const bf16_t *)X, t, N, batch,
                         seed + 0x632BE59BD9B4E019ull * (unsigned long long)(k + 1), s->Hs[0], s->tb);
      s->forward(batch);
      backward_full(s, batch);
      HIPCHK(hipMemcpyAsync(dl + k, s->loss, sizeof(double), hipMemcpyDeviceToDevice, s->st));
      const float frac = nsteps > 1 ? (float)k / (float)(nsteps - 1) : 0.0f;
      const float lrk = lr * (0.01f + 0.99f * 0.5f * (1.0f + cosf(3.14159265f * frac)));
      adam(s, lrk);
    }
    HIPCHK(hipGetLastError());
    if (losses) HIPCHK(hipMemcpyAsync(losses, dl, nsteps * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

// n_take rows drawn uniformly (with replacement) from device rows (x_src [n_src x 64],
// t_src [n_src]) into x_dst/t_dst: the per-configuration sample of the training set.
extern "C" int nrpyell_mlp_sample_rows(const void *x_src, const float *t_src, long n_src, long n_take,
                                       unsigned long long seed, void *x_dst, float *t_dst) {
  return api_call([&]() -> int {
    if (n_src <= 0 || n_take <= 0 || n_take > (1L << 30)) {
      set_last_error("sample_rows: bad sizes");
      return NRPYELL_EINVAL;
    }
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)n_take), dim3(64), 0, 0, (const bf16_t *)x_src, t_src, n_src,
                       (int)n_take, seed, (bf16_t *)x_dst, t_dst);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NRPYELL_OK;
  });
}

extern "C" int nrpyell_mlp_adam_step(nrpyell_mlp *s, float lr) {
  return api_call([&]() -> int {
    adam(s, lr);
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

static FeatureParams feature_params(const paramstruct *p) {
  FeatureParams f;
  f.AMAX = p->AMAX;
  f.w = p->SINHWAA;
  f.b = p->bScale;
  f.D = exp(1.0 / p->SINHWAA) - exp(-1.0 / p->SINHWAA);
  f.m[0] = p->bare_mass_0;
  f.m[1] = p->bare_mass_1;
  const double pos[2][3] = {{p->puncture_0_x, p->puncture_0_y, p->puncture_0_z},
                            {p->puncture_1_x, p->puncture_1_y, p->puncture_1_z}};
  const double P[2][3] = {{p->P0_x, p->P0_y, p->P0_z}, {p->P1_x, p->P1_y, p->P1_z}};
  const double S[2][3] = {{p->S0_x, p->S0_y, p->S0_z}, {p->S1_x, p->S1_y, p->S1_z}};
  memcpy(f.pos, pos, sizeof(pos));
  memcpy(f.P, P, sizeof(P));
  memcpy(f.S, S, sizeof(S));
  return f;
}

// Number of feature rows of a grid (interior + pure outer points), rounded up to 64.
extern "C" long nrpyell_mlp_feature_rows(const griddata_struct *gd) {
  const GridDims g = dims_from_params(&gd->params);
  const long n = (long)(g.G0 - NRPYELL_NG) * g.N1 * g.N2;
  return (n + 63) / 64 * 64;
}

// Features (device BF16 [rows x 32], padding rows zero), per-row u scale and grid index.
extern "C" int nrpyell_mlp_features(const griddata_struct *gd, void *X, double *scale, long *idx) {
  return api_call([&]() -> int {
    const GridDims g = dims_from_params(&gd->params);
    const long n = (long)(g.G0 - NRPYELL_NG) * g.N1 * g.N2, rows = (n + 63) / 64 * 64;
    DevBuf<double> b0(g.G0), b1(g.G1), b2(g.G2);
    double *x0 = b0.p, *x1 = b1.p, *x2 = b2.p;
    HIPCHK(hipMemcpy(x0, gd->xx[0], g.G0 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x1, gd->xx[1], g.G1 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x2, gd->xx[2], g.G2 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(X, 0, rows * MLP_K0 * sizeof(bf16_t)));
    hipLaunchKernelGGL(features_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, g, feature_params(&gd->params), x0, x1,
                       x2, (bf16_t *)X, scale, idx);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NRPYELL_OK;
  });
}

// Training targets t = u/scale from the solved y_n of a griddata at the feature rows.
extern "C" int nrpyell_mlp_targets(const griddata_struct *gd, const double *scale, const long *idx, float *t) {
  return api_call([&]() -> int {
    const GridDims g = dims_from_params(&gd->params);
    const long n = (long)(g.G0 - NRPYELL_NG) * g.N1 * g.N2;
    hipLaunchKernelGGL(target_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, gd->gridfuncs.y_n_gfs, scale, idx, t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NRPYELL_OK;
  });
}

// Seed griddata's y_n from the network: u = e^t scale at interior and pure outer points,
// smoothed by smooth_passes binomial passes (each along i0, i1, i2), v = eta u (the steady
// state of du/dt = v - eta u), parity images written in every ghost point. k_odd_gfs is the
// scratch of the smoothing.
extern "C" int nrpyell_mlp_seed(nrpyell_mlp *s, griddata_struct *gd, int smooth_passes) {
  return api_call([&]() -> int {
    if (smooth_passes < 0 || smooth_passes > 64) {
      set_last_error("mlp_seed: smooth_passes must be in [0, 64]");
      return NRPYELL_EINVAL;
    }
    const GridDims g = dims_from_params(&gd->params);
    const long n = (long)(g.G0 - NRPYELL_NG) * g.N1 * g.N2, rows = (n + 63) / 64 * 64;
    DevBuf<bf16_t> Xb(rows * MLP_K0);
    DevBuf<double> scaleb(rows);
    DevBuf<long> idxb(rows);
    DevBuf<float> yb(rows);
    bf16_t *X = Xb.p;
    double *scale = scaleb.p;
    long *idx = idxb.p;
    float *y = yb.p;
    int rc = nrpyell_mlp_features(gd, X, scale, idx);
    if (rc != NRPYELL_OK) return rc;
    const int chunk = 1 << 18;  // rows per forward pass (multiple of 64)
    for (long r0 = 0; r0 < rows; r0 += chunk) {
      const int M = (int)((rows - r0) < chunk ? (rows - r0) : chunk);
      if ((rc = nrpyell_mlp_forward(s, X + r0 * MLP_K0, M, y + r0)) != NRPYELL_OK) return rc;
    }
    double *yn = gd->gridfuncs.y_n_gfs, *tmp = gd->gridfuncs.k_odd_gfs;
    hipLaunchKernelGGL(seed_state_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, g, n, y, scale, idx,
                       gd->params.eta_damping, yn);
    const dim3 sg((g.G0 - NRPYELL_NG + 63) / 64, g.N1, g.N2);
    for (int k = 0; k < smooth_passes; k++) {
      hipLaunchKernelGGL(smooth_kernel<0>, sg, dim3(64), 0, 0, g, yn, tmp);
      hipLaunchKernelGGL(smooth_kernel<1>, sg, dim3(64), 0, 0, g, tmp, yn);
      hipLaunchKernelGGL(smooth_kernel<2>, sg, dim3(64), 0, 0, g, yn, tmp);
      HIPCHK(hipMemcpyAsync(yn, tmp, sizeof(double) * g.ntot, hipMemcpyDeviceToDevice, 0));
    }
    if (smooth_passes > 0) {
      launch_fill_ghost_planes<double>(g, yn, 1, 0);
      hipLaunchKernelGGL(v_from_u_kernel, dim3((g.ntot + 255) / 256), dim3(256), 0, 0, g.ntot,
                         gd->params.eta_damping, yn);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NRPYELL_OK;
  });
}

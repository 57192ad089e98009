// nrpyell_mlp.hip -- the learned initial guess: a coordinate MLP u(x; punctures), trained on
// solutions this solver produced, run on hand-written BF16 MFMA kernels (CDNA4
// v_mfma_f32_32x32x16_bf16) with FP32 accumulation. No reference counterpart.
//
// Network: 23 input features (zero-padded to K0 = 64) -> 5 x [Linear(256) + SiLU] -> Linear(1).
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

// splitmix64 of (seed, m): the dataset row of minibatch row m (uniform, with replacement)
__device__ __forceinline__ long draw_row(unsigned long long seed, int m, long N) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(m + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (long)(z % (unsigned long long)N);
}

// ---------------------------------------------------------------------------------------
// Training: one minibatch step is three launches.
//
// 1. mlp_train_rows_kernel: a block takes TR_R rows through the forward pass (activations in
//    LDS, ping-pong), the output layer and the loss, then back through the data-gradient chain
//    dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}). Every product is row-local, so nothing crosses
//    blocks. silu'(z) of each layer stays in registers: the lane that produces z in the forward
//    epilogue holds the same (row, column) entries of dZ in the backward one. H_l (l = 0..5)
//    and dZ_l (l = 1..5) leave the block transposed ([width x M], rows contiguous), which is the
//    layout the weight gradients consume as MFMA operands.
// 2. mlp_wgrad_kernel: dW_l = dZ_l^T H_{l-1}, db_l = column sums of dZ_l, and the output layer's
//    gradients, over TR_SPLITS row ranges; each range writes its own partial (no atomics).
// 3. mlp_adam_kernel: g = the partials summed in range order (deterministic), Adam, and the BF16
//    weight copies (W_l and, for l >= 2, W_l transposed, the data-gradient chain's operand).
//
// Arithmetic per entry matches the network definition: FP32 accumulation of BF16 products in
// ascending inner index, z = acc + b, h = bf16(silu(z)), y = sum_k h_k w_k + b (FP32),
// dy = 2 (y - t) / M, dZ_5 = bf16(dy w_out silu'(z_5)), dZ_{l-1} = bf16(acc silu'(z_{l-1})).
// ---------------------------------------------------------------------------------------
constexpr int TR_R = 32;        // rows per block
constexpr int TR_WAVES = 8;     // wave w owns columns 32w .. 32w+31 of every 256-wide layer
constexpr int TR_LD = MLP_W + 8;
constexpr int TR_SPLITS = 32;   // row ranges of the weight-gradient reduction
constexpr int WG_GROUPS = MLP_HIDDEN * TR_SPLITS;  // (layer, row range) groups of the weight gradients
constexpr int TR_LDS_DSL = 3;         // layers whose silu'(z) the row kernel keeps in LDS
constexpr int WG_G = 2;               // inner steps per load group of the weight-gradient kernel

// Operand layout of a matrix X[rows][inner] (rows a multiple of 32, inner a multiple of 16):
// lane (li, lj) of a wave feeds the 32x32x16 MFMA for rows 32 t .. 32 t + 31 and inner step q
// with X[32 t + li][16 q + 8 lj .. 16 q + 8 lj + 7]; the layout keeps those 8 elements at
// ((t * (inner / 16) + q) * 64 + 32 lj + li) * 8, so a wave's operand load is 1 KB of
// consecutive bytes (a row-major load would touch 32 rows, 32 bytes each).
__device__ __forceinline__ long swz(int row, int col, int inner16) {
  return ((long)((row >> 5) * inner16 + (col >> 4)) * 64 + ((col >> 3) & 1) * 32 + (row & 31)) * 8 + (col & 7);
}

// Where the BF16 weight copies live: pw[k] = bf16(p[k]) (row-major, for inference) and, in
// pws, W_l (rows n, inner k; l = 1..5) and W_l^T (rows k, inner n; l = 2..5) in the operand layout.
struct WeightCopies {
  long w_off[MLP_HIDDEN + 1];  // W_l in the FP32 parameters
  long fs[MLP_HIDDEN + 1];     // W_l in pws
  long ts[MLP_HIDDEN + 1];     // W_l^T in pws
};

struct TrainArgs {
  const bf16_t *X;  // dataset features [N x 64]
  const float *t;   // dataset targets [N]
  long N;
  unsigned long long seed;
  int gather;  // 1: minibatch row m is dataset row draw_row(seed, m, N); 0: row m
  int M;       // minibatch rows (multiple of 64)
  const bf16_t *pws;  // weights in the operand layout (WeightCopies)
  const float *p;
  MlpOffsets off;
  WeightCopies wc;
  bf16_t *Ht[MLP_HIDDEN + 1];      // H_l^T [width_l x M], l = 0..5
  bf16_t *dZt[MLP_HIDDEN + 1];     // dZ_l^T [256 x M], l = 1..5
  float *dy;                       // [M]
  double *loss;                    // += sum over the block's rows of (y - t)^2 / M
};

// rows r0 .. r0+31 of an LDS tile (width WD) -> the transposed global [WD x M] in the operand
// layout (rows = the WD columns, inner = minibatch rows)
template <int WD>
__device__ __forceinline__ void store_transposed(const bf16_t *buf, bf16_t *__restrict__ out, int r0, int M, int tid) {
  constexpr int PER = TR_R * WD / (64 * TR_WAVES);  // rows per thread: 16 (WD 256) or 4 (WD 64)
  const int c = tid % WD, q = tid / WD;
  const int m = r0 + PER * q, M16 = M / 16;
  bf16_t v[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) v[i] = buf[(PER * q + i) * TR_LD + c];
  if (PER == 16) {
    bf16x8 lo, hi;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      lo[i] = v[i];
      hi[i] = v[8 + i];
    }
    *(bf16x8 *)(out + swz(c, m, M16)) = lo;
    *(bf16x8 *)(out + swz(c, m + 8, M16)) = hi;
  } else {
    bf16_t *o = out + swz(c, m, M16);
#pragma unroll
    for (int i = 0; i < PER; i++) o[i] = v[i];
  }
}

__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoidf_(z);
  return s * (1.0f + z * (1.0f - s));
}

__global__ void __launch_bounds__(512) mlp_train_rows_kernel(TrainArgs a) {
  __shared__ __attribute__((aligned(16))) bf16_t Hb[2][TR_R * TR_LD];
  __shared__ float s_part[TR_WAVES][TR_R];
  __shared__ float s_t[TR_R], s_dy[TR_R];
  __shared__ long s_row[TR_R];
  // silu'(z) of layers 1..TR_LDS_DSL, each thread's 16 entries at [layer][r][thread] (read back by
  // the same thread); layers above stay in registers
  __shared__ float s_dsl[TR_LDS_DSL][16][64 * TR_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 31, lj = lane >> 5;
  const int r0 = blockIdx.x * TR_R;
  // this lane's column of every 256-wide layer; the eight 32-column slices are rotated by the
  // block index so that concurrent blocks read different weight rows (no change to any sum)
  const int ct = (wave + blockIdx.x) & (TR_WAVES - 1);
  const int c = ct * 32 + li;
  if (tid < TR_R) {
    const long row = a.gather ? draw_row(a.seed, r0 + tid, a.N) : (long)(r0 + tid);
    s_row[tid] = row;
    s_t[tid] = a.t[row];
  }
  __syncthreads();
  if (tid < TR_R * 8) {  // 32 rows x 64 BF16 features = 256 chunks of 16 B
    const int row = tid >> 3, ch = tid & 7;
    *(bf16x8 *)(&Hb[0][row * TR_LD + 8 * ch]) = *(const bf16x8 *)(a.X + s_row[row] * MLP_K0 + 8 * ch);
  }
  __syncthreads();
  store_transposed<MLP_K0>(Hb[0], a.Ht[0], r0, a.M, tid);
  // accumulator r of lane (li, lj): row 8*(r/4) + 4*lj + r%4 of the block, column c
  float dsl[MLP_HIDDEN][16];
  // the weight operands of the next product are loaded (all of them, 16 B per lane per
  // 16-wide inner step) as soon as the current product's MFMAs are issued, so their latency
  // overlaps the epilogue, the barrier and the transposed store
  bf16x8 bv[MLP_W / 16];
#pragma unroll
  for (int q = 0; q < MLP_K0 / 16; q++)
    bv[q] = *(const bf16x8 *)(a.pws + a.wc.fs[1] + ((long)ct * (MLP_K0 / 16) + q) * 512 + 8 * lane);
  int cur = 0;
#pragma unroll
  for (int l = 1; l <= MLP_HIDDEN; l++) {
    const int Kl = l == 1 ? MLP_K0 : MLP_W;
    const bf16_t *Hin = Hb[cur] + li * TR_LD + 8 * lj;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
    for (int q = 0; q < Kl / 16; q++)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*(const bf16x8 *)(Hin + 16 * q), bv[q], acc, 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    // next: W_{l+1}, or W_5^T for the first step of the data-gradient chain
    const bf16_t *Wn = a.pws + (l < MLP_HIDDEN ? a.wc.fs[l + 1] : a.wc.ts[MLP_HIDDEN]) + (long)ct * 16 * 512 + 8 * lane;
#pragma unroll
    for (int q = 0; q < MLP_W / 16; q++) bv[q] = *(const bf16x8 *)(Wn + 512 * q);
    bf16_t *Hout = Hb[cur ^ 1];
    const float bc = a.p[a.off.b[l] + c];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = 8 * (r >> 2) + 4 * lj + (r & 3);
      const float z = acc[r] + bc;
      if (l <= TR_LDS_DSL)
        s_dsl[l - 1][r][tid] = dsilu(z);
      else
        dsl[l - 1][r] = dsilu(z);
      Hout[row * TR_LD + c] = to_bf16(z * sigmoidf_(z));
    }
    __syncthreads();
    cur ^= 1;
    store_transposed<MLP_W>(Hb[cur], a.Ht[l], r0, a.M, tid);
  }
  // output layer: wave w sums columns 32w .. 32w+31 of every row, then the 8 partials in a fixed order
  const float *wout = a.p + a.off.w[MLP_HIDDEN + 1];
  if (lane < TR_R) {
    const bf16_t *h = Hb[cur] + lane * TR_LD + wave * 32;
    float acc = 0.0f;
#pragma unroll 8
    for (int k = 0; k < 32; k++) acc += (float)h[k] * wout[wave * 32 + k];
    s_part[wave][lane] = acc;
  }
  __syncthreads();
  if (tid < 64) {
    double l2 = 0.0;
    if (tid < TR_R) {
      const float y = (((s_part[0][tid] + s_part[1][tid]) + (s_part[2][tid] + s_part[3][tid])) +
                       ((s_part[4][tid] + s_part[5][tid]) + (s_part[6][tid] + s_part[7][tid]))) +
                      a.p[a.off.b[MLP_HIDDEN + 1]];
      const float diff = y - s_t[tid];
      const float dy = 2.0f * diff / (float)a.M;
      s_dy[tid] = dy;
      a.dy[r0 + tid] = dy;
      l2 = (double)diff * (double)diff / (double)a.M;
    }
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_down(l2, o, 64);
    if (tid == 0) atomicAdd(a.loss, l2);
  }
  __syncthreads();
  // dZ_5 into the other buffer (H_4 there is no longer read)
  {
    bf16_t *D = Hb[cur ^ 1];
    const float wc_ = wout[c];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = 8 * (r >> 2) + 4 * lj + (r & 3);
      D[row * TR_LD + c] = to_bf16(s_dy[row] * wc_ * dsl[MLP_HIDDEN - 1][r]);
    }
  }
  __syncthreads();
  cur ^= 1;
  store_transposed<MLP_W>(Hb[cur], a.dZt[MLP_HIDDEN], r0, a.M, tid);
  // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}), l = 5..2: A = dZ_l (LDS), B(k, n) = W_l^T[k][n]
#pragma unroll
  for (int l = MLP_HIDDEN; l >= 2; l--) {
    const bf16_t *Din = Hb[cur] + li * TR_LD + 8 * lj;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
    for (int q = 0; q < MLP_W / 16; q++)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*(const bf16x8 *)(Din + 16 * q), bv[q], acc, 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    if (l > 2) {
      const bf16_t *WT = a.pws + a.wc.ts[l - 1] + (long)ct * 16 * 512 + 8 * lane;
#pragma unroll
      for (int q = 0; q < MLP_W / 16; q++) bv[q] = *(const bf16x8 *)(WT + 512 * q);
    }
    bf16_t *Dout = Hb[cur ^ 1];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = 8 * (r >> 2) + 4 * lj + (r & 3);
      Dout[row * TR_LD + c] = to_bf16(acc[r] * (l - 1 <= TR_LDS_DSL ? s_dsl[l - 2][r][tid] : dsl[l - 2][r]));
    }
    __syncthreads();
    cur ^= 1;
    store_transposed<MLP_W>(Hb[cur], a.dZt[l - 1], r0, a.M, tid);
  }
}

struct WgradArgs {
  const bf16_t *Ht[MLP_HIDDEN + 1];
  const bf16_t *dZt[MLP_HIDDEN + 1];
  const float *dy;
  int M, chunk;  // rows per range (multiple of 16)
  MlpOffsets off;
  long n_params;
  float *part;   // [TR_SPLITS x n_params]
};

// One wave's (32 NT) x (32 KT) block of C[n][k] = sum_{m in [m0, m1)} A[n][m] B[k][m] for the
// rows n = 32 ta .., k = 32 tb .. of two transposed minibatch arrays in the operand layout (inner
// = the M minibatch rows), loaded straight from global as MFMA operands (1 KB per wave and load)
// in groups of WG_G 16-row inner steps, double-buffered (the loads of group g+1 are in flight
// while group g's MFMAs run). C goes to G (row stride ldg); with gb != null the row sums of A
// (bias gradient) go to gb.
template <int NT, int KT>
__device__ __forceinline__ void wgrad_block(const bf16_t *__restrict__ A, int ta, const bf16_t *__restrict__ B, int tb,
                                            int M, int m0, int m1, float *__restrict__ G, int ldg,
                                            float *__restrict__ gb) {
  const int lane = threadIdx.x & 63, li = lane & 31, lj = lane >> 5;
  const long M16 = M / 16;
  const bf16_t *Ap[NT], *Bp[KT];
#pragma unroll
  for (int x = 0; x < NT; x++) Ap[x] = A + ((ta + x) * M16 * 64 + lane) * 8;
#pragma unroll
  for (int y = 0; y < KT; y++) Bp[y] = B + ((tb + y) * M16 * 64 + lane) * 8;
  floatx16 acc[NT][KT];
#pragma unroll
  for (int x = 0; x < NT; x++)
#pragma unroll
    for (int y = 0; y < KT; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.0f;
  float bsum[NT];
#pragma unroll
  for (int x = 0; x < NT; x++) bsum[x] = 0.0f;
  const bool bias = gb != nullptr;
  const int nsteps = m1 > m0 ? (m1 - m0) / 16 : 0;
  const int ngroups = (nsteps + WG_G - 1) / WG_G;
  bf16x8 a0[WG_G][NT], b0[WG_G][KT], a1[WG_G][NT], b1[WG_G][KT];
  auto load = [&](bf16x8 (*av)[NT], bf16x8 (*bw)[KT], int g) {
#pragma unroll
    for (int q = 0; q < WG_G; q++)
      if (g * WG_G + q < nsteps) {
        const long o = (long)(m0 / 16 + g * WG_G + q) * 512;  // inner step m / 16
#pragma unroll
        for (int x = 0; x < NT; x++) av[q][x] = *(const bf16x8 *)(Ap[x] + o);
#pragma unroll
        for (int y = 0; y < KT; y++) bw[q][y] = *(const bf16x8 *)(Bp[y] + o);
      }
  };
  auto mma = [&](bf16x8 (*av)[NT], bf16x8 (*bw)[KT], int g) {
#pragma unroll
    for (int q = 0; q < WG_G; q++)
      if (g * WG_G + q < nsteps) {
#pragma unroll
        for (int x = 0; x < NT; x++)
#pragma unroll
          for (int y = 0; y < KT; y++)
            acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[q][x], bw[q][y], acc[x][y], 0, 0, 0);
        if (bias) {
#pragma unroll
          for (int x = 0; x < NT; x++)
#pragma unroll
            for (int i = 0; i < 8; i++) bsum[x] += (float)av[q][x][i];
        }
      }
  };
  if (ngroups > 0) load(a0, b0, 0);
  for (int g = 0; g < ngroups; g += 2) {
    if (g + 1 < ngroups) load(a1, b1, g + 1);
    __builtin_amdgcn_sched_barrier(0);
    mma(a0, b0, g);
    __builtin_amdgcn_sched_barrier(0);
    if (g + 2 < ngroups) load(a0, b0, g + 2);
    __builtin_amdgcn_sched_barrier(0);
    if (g + 1 < ngroups) mma(a1, b1, g + 1);
    __builtin_amdgcn_sched_barrier(0);
  }
  // accumulator r of lane (li, lj) is row 8*(r/4) + 4*lj + r%4, column li of its 32 x 32 tile
#pragma unroll
  for (int x = 0; x < NT; x++)
#pragma unroll
    for (int y = 0; y < KT; y++)
#pragma unroll
      for (int r = 0; r < 16; r++)
        G[(long)(32 * x + 8 * (r >> 2) + 4 * lj + (r & 3)) * ldg + 32 * y + li] = acc[x][y][r];
  if (bias) {
#pragma unroll
    for (int x = 0; x < NT; x++) {
      const float t = bsum[x] + __shfl_xor(bsum[x], 32, 64);
      if (lj == 0) gb[32 * x + li] = t;
    }
  }
}

// Weight gradients over row ranges. Blocks 0 .. 4*WG_GROUPS-1: group (layer l, range s) has four
// blocks nt = 0..3, each the 64 rows 64 nt .. 64 nt + 63 of dW_l (all K_l columns): for l >= 2
// wave w takes columns 64 w .. 64 w + 63 (2 x 2 MFMA tiles), for l = 1 wave w the 32 x 32
// quarter (w / 2, w % 2) of the 64 x 64 block. Block b = 32 (g / 8) + 8 nt + g % 8 for group g,
// so the four blocks of a group share b mod 8, i.e. one XCD (workgroups are dealt to the XCDs
// round-robin) and one L2, where the group's dZ_l and H_{l-1} row slices are reused. The last
// TR_SPLITS blocks: the output layer's gradients, one range each.
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(WgradArgs a) {
  const int tid = threadIdx.x, wave = tid >> 6;
  const int b = blockIdx.x;
  if (b >= 4 * WG_GROUPS) {
    const int s = b - 4 * WG_GROUPS;
    const int m0 = s * a.chunk, m1 = min(a.M, m0 + a.chunk);
    float *P = a.part + (long)s * a.n_params;
    // gw[k] = sum_m dy[m] H_5[m][k], gb = sum_m dy[m]; thread k reads row k of H_5^T (operand layout)
    const bf16_t *h = a.Ht[MLP_HIDDEN] + (((long)(tid >> 5) * (a.M / 16)) * 64 + (tid & 31)) * 8;
    float acc = 0.0f, accb = 0.0f;
    for (int m = m0; m < m1; m += 8) {
      const bf16x8 hv = *(const bf16x8 *)(h + ((long)(m >> 4) * 64 + ((m >> 3) & 1) * 32) * 8);
#pragma unroll
      for (int i = 0; i < 8; i++) {
        acc += a.dy[m + i] * (float)hv[i];
        accb += a.dy[m + i];
      }
    }
    P[a.off.w[MLP_HIDDEN + 1] + tid] = acc;
    if (tid == 0) P[a.off.b[MLP_HIDDEN + 1]] = accb;
    return;
  }
  const int q = b >> 5, rem = b & 31, nt = rem >> 3, g = q * 8 + (rem & 7);
  const int l = 1 + g / TR_SPLITS, s = g % TR_SPLITS;
  const int m0 = s * a.chunk, m1 = min(a.M, m0 + a.chunk);
  float *P = a.part + (long)s * a.n_params;
  if (l == 1) {
    const int n0 = nt * 64 + (wave >> 1) * 32, k0 = (wave & 1) * 32;
    wgrad_block<1, 1>(a.dZt[1], n0 / 32, a.Ht[0], k0 / 32, a.M, m0, m1,
                      P + a.off.w[1] + (long)n0 * MLP_K0 + k0, MLP_K0,
                      (wave & 1) == 0 ? P + a.off.b[1] + n0 : nullptr);
  } else {
    const int n0 = nt * 64, k0 = wave * 64;
    wgrad_block<2, 2>(a.dZt[l], n0 / 32, a.Ht[l - 1], k0 / 32, a.M, m0, m1,
                      P + a.off.w[l] + (long)n0 * MLP_W + k0, MLP_W, wave == 0 ? P + a.off.b[l] + n0 : nullptr);
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

// the BF16 copies of parameter k (see WeightCopies)
__device__ __forceinline__ void store_weight(const WeightCopies &wc, long k, float v, bf16_t *__restrict__ pw,
                                             bf16_t *__restrict__ pws) {
  const bf16_t b = to_bf16(v);
  pw[k] = b;
#pragma unroll
  for (int l = 1; l <= MLP_HIDDEN; l++) {
    const int Kl = l == 1 ? MLP_K0 : MLP_W;
    const long d = k - wc.w_off[l];
    if (d >= 0 && d < (long)MLP_W * Kl) {
      const int n = (int)(d / Kl), kk = (int)(d % Kl);
      pws[wc.fs[l] + swz(n, kk, Kl / 16)] = b;
      if (l >= 2) pws[wc.ts[l] + swz(kk, n, MLP_W / 16)] = b;
    }
  }
}

__global__ void refresh_weights_kernel(long n, const float *__restrict__ p, WeightCopies wc, bf16_t *__restrict__ pw,
                                       bf16_t *__restrict__ pws) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) store_weight(wc, k, p[k], pw, pws);
}

// g = sum of the TR_SPLITS partials in range order; with update, an Adam step on p and the
// BF16 copies refreshed.
__global__ void mlp_adam_kernel(long n, const float *__restrict__ part, float *__restrict__ g, int update,
                                float *__restrict__ p, float *__restrict__ mo, float *__restrict__ ve, float lr,
                                float b1, float b2, float eps, float c1, float c2, WeightCopies wc,
                                bf16_t *__restrict__ pw, bf16_t *__restrict__ pws) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float gk = 0.0f;
#pragma unroll
  for (int s = 0; s < TR_SPLITS; s++) gk += part[(long)s * n + k];
  g[k] = gk;
  if (!update) return;
  const float m = b1 * mo[k] + (1.0f - b1) * gk;
  const float v = b2 * ve[k] + (1.0f - b2) * gk * gk;
  mo[k] = m;
  ve[k] = v;
  const float pk = p[k] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  p[k] = pk;
  store_weight(wc, k, pk, pw, pws);
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
  float *part = nullptr;  // per-range gradient partials [TR_SPLITS x n_params]
  bf16_t *pw = nullptr;   // BF16 copy of all parameters (weights read by the MFMA kernels)
  bf16_t *pws = nullptr;  // BF16 W_l and W_l^T in the MFMA operand layout (WeightCopies)
  long adam_t = 0;
  int cap = 0;                       // minibatch arrays allocated for this many rows
  bf16_t *Ht[MLP_HIDDEN + 1] = {};   // H_l^T [width_l x M], l = 0..5
  bf16_t *dZt[MLP_HIDDEN + 1] = {};  // dZ_l^T [256 x M], l = 1..5
  float *dy = nullptr;               // per-row output gradient
  double *loss = nullptr;
  // seeding workspace (features, per-row scale and point index, network output), grown on demand
  // and kept between seeds so a seed allocates nothing after the first one on a grid this large
  long seed_cap = 0;
  bf16_t *seedX = nullptr;
  double *seed_scale = nullptr;
  long *seed_idx = nullptr;
  float *seed_y = nullptr;
  hipStream_t st = nullptr;
  ~nrpyell_mlp() {
    (void)hipFree(p); (void)hipFree(g); (void)hipFree(m); (void)hipFree(v); (void)hipFree(part);
    (void)hipFree(pw); (void)hipFree(pws); (void)hipFree(loss);
    release_seed();
    release();
    if (st) (void)hipStreamDestroy(st);
  }
  int K(int l) const { return l == 1 ? MLP_K0 : MLP_W; }
  void release() {
    for (auto *&h : Ht) { (void)hipFree(h); h = nullptr; }
    for (auto *&d : dZt) { (void)hipFree(d); d = nullptr; }
    (void)hipFree(dy);
    dy = nullptr;
    cap = 0;
  }
  void release_seed() {
    (void)hipFree(seedX); (void)hipFree(seed_scale); (void)hipFree(seed_idx); (void)hipFree(seed_y);
    seedX = nullptr; seed_scale = nullptr; seed_idx = nullptr; seed_y = nullptr;
    seed_cap = 0;
  }
  void ensure_seed(long rows) {
    if (rows <= seed_cap) return;
    release_seed();  // a failed allocation below leaves seed_cap 0: the next seed allocates again
    seedX = mlp_alloc<bf16_t>((size_t)rows * MLP_K0);
    seed_scale = mlp_alloc<double>((size_t)rows);
    seed_idx = mlp_alloc<long>((size_t)rows);
    seed_y = mlp_alloc<float>((size_t)rows);
    seed_cap = rows;
  }
  void ensure(int M) {
    if (M <= cap) return;
    // release first; a failed allocation below (HipFailure) leaves cap 0 and null or valid
    // pointers only, so the next call allocates again
    release();
    Ht[0] = mlp_alloc<bf16_t>((size_t)M * MLP_K0);
    for (int l = 1; l <= MLP_HIDDEN; l++) {
      Ht[l] = mlp_alloc<bf16_t>((size_t)M * MLP_W);
      dZt[l] = mlp_alloc<bf16_t>((size_t)M * MLP_W);
    }
    dy = mlp_alloc<float>(M);
    cap = M;
  }
  MlpOffsets offsets() const {
    MlpOffsets off;
    for (int l = 1; l <= MLP_HIDDEN + 1; l++) {
      off.w[l] = w_off[l];
      off.b[l] = b_off[l];
    }
    return off;
  }
  WeightCopies copies() const {
    WeightCopies wc;
    long o = 0;
    for (int l = 1; l <= MLP_HIDDEN; l++) {
      wc.w_off[l] = w_off[l];
      wc.fs[l] = o;
      o += (long)MLP_W * K(l);
    }
    for (int l = 2; l <= MLP_HIDDEN; l++) {
      wc.ts[l] = o;
      o += (long)MLP_W * MLP_W;
    }
    wc.w_off[0] = wc.fs[0] = wc.ts[0] = wc.ts[1] = 0;
    return wc;
  }
  static long pws_size() { return (long)MLP_W * MLP_K0 + (2L * MLP_HIDDEN - 2) * MLP_W * MLP_W; }
  void refresh_bf16() {
    hipLaunchKernelGGL(refresh_weights_kernel, dim3((n_params + 255) / 256), dim3(256), 0, st, n_params, p, copies(),
                       pw, pws);
  }
  // forward, loss (added to *loss_out) and gradient partials of M minibatch rows (M a multiple
  // of 64): rows drawn from the dataset (gather) or its first M rows
  void loss_partials(const bf16_t *X, const float *t, long N, unsigned long long seed, int gather, int M,
                     double *loss_out) {
    ensure(M);
    TrainArgs a;
    a.X = X;
    a.t = t;
    a.N = N;
    a.seed = seed;
    a.gather = gather;
    a.M = M;
    a.pws = pws;
    a.wc = copies();
    a.p = p;
    a.off = offsets();
    for (int l = 0; l <= MLP_HIDDEN; l++) {
      a.Ht[l] = Ht[l];
      a.dZt[l] = dZt[l];
    }
    a.dy = dy;
    a.loss = loss_out;
    hipLaunchKernelGGL(mlp_train_rows_kernel, dim3(M / TR_R), dim3(64 * TR_WAVES), 0, st, a);
    WgradArgs w;
    for (int l = 0; l <= MLP_HIDDEN; l++) {
      w.Ht[l] = Ht[l];
      w.dZt[l] = dZt[l];
    }
    w.dy = dy;
    w.M = M;
    w.chunk = ((M + TR_SPLITS - 1) / TR_SPLITS + 15) / 16 * 16;
    w.off = a.off;
    w.n_params = n_params;
    w.part = part;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(4 * WG_GROUPS + TR_SPLITS), dim3(256), 0, st, w);
  }
  // g = the summed partials; with update, one Adam step at learning rate lr
  void reduce(int update, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f, c2 = 1.0f;
    if (update) {
      adam_t++;
      c1 = 1.0f - powf(b1, (float)adam_t);
      c2 = 1.0f - powf(b2, (float)adam_t);
    }
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((n_params + 255) / 256), dim3(256), 0, st, n_params, part, g, update, p,
                       m, v, lr, b1, b2, 1e-8f, c1, c2, copies(), pw, pws);
  }
};

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
    s->pws = mlp_alloc<bf16_t>((size_t)nrpyell_mlp::pws_size());
    s->part = mlp_alloc<float>((size_t)TR_SPLITS * off);
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

// Forward pass on M rows of device BF16 features [M x 64]; y (device, FP32) gets M outputs.
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
// Loss (mean squared error) and its gradient for M rows (device BF16 features [M x 64], FP32
// targets); grads copied to host when non-null.
extern "C" int nrpyell_mlp_loss_grad(nrpyell_mlp *s, const void *x, const float *t, int M, double *loss,
                                     float *grads_host) {
  return api_call([&]() -> int {
    if (check_rows(M)) return NRPYELL_EINVAL;
    HIPCHK(hipMemsetAsync(s->loss, 0, sizeof(double), s->st));
    s->loss_partials((const bf16_t *)x, t, M, 0, 0, M, s->loss);
    s->reduce(0, 0.0f);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(loss, s->loss, sizeof(double), hipMemcpyDeviceToHost, s->st));
    if (grads_host) HIPCHK(hipMemcpyAsync(grads_host, s->g, s->n_params * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return NRPYELL_OK;
  });
}

// nsteps of Adam on minibatches of `batch` rows drawn uniformly from a device dataset
// (features BF16 [N x 64], targets FP32 [N]); cosine learning-rate decay from lr to lr/100.
// losses (host, nsteps entries, may be null) receives each step's minibatch loss. Three
// launches per step (rows, weight gradients, Adam); the seed of step k's draw is
// seed + 0x632BE59BD9B4E019 (k + 1).
extern "C" int nrpyell_mlp_train(nrpyell_mlp *s, const void *X, const float *t, long N, int batch, int nsteps,
                                 float lr, unsigned long long seed, double *losses) {
  return api_call([&]() -> int {
    if (check_rows(batch) || N <= 0 || nsteps < 0) return NRPYELL_EINVAL;
    DevBuf<double> dlb(nsteps > 0 ? nsteps : 1);
    double *dl = dlb.p;
    HIPCHK(hipMemsetAsync(dl, 0, (nsteps > 0 ? nsteps : 1) * sizeof(double), s->st));
    for (int k = 0; k < nsteps; k++) {
      s->loss_partials((const bf16_t *)X, t, N, seed + 0x632BE59BD9B4E019ull * (unsigned long long)(k + 1), 1, batch,
                       dl + k);
      const float frac = nsteps > 1 ? (float)k / (float)(nsteps - 1) : 0.0f;
      const float lrk = lr * (0.01f + 0.99f * 0.5f * (1.0f + cosf(3.14159265f * frac)));
      s->reduce(1, lrk);
    }
    HIPCHK(hipGetLastError());
    if (losses && nsteps > 0)
      HIPCHK(hipMemcpyAsync(losses, dl, nsteps * sizeof(double), hipMemcpyDeviceToHost, s->st));
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

// One Adam step with the gradient of the last nrpyell_mlp_loss_grad call (its partials).
extern "C" int nrpyell_mlp_adam_step(nrpyell_mlp *s, float lr) {
  return api_call([&]() -> int {
    s->reduce(1, lr);
    HIPCHK(hipGetLastError());
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

// Features (device BF16 [rows x 64], padding rows zero), per-row u scale and grid index.
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
    s->ensure_seed(rows);
    bf16_t *X = s->seedX;
    double *scale = s->seed_scale;
    long *idx = s->seed_idx;
    float *y = s->seed_y;
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

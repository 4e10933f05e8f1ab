// costmodel.hip - the learned root-cost model: a BF16 MLP 8 -> 64 -> 64 -> 2
// with ReLU, evaluated and trained by hand-written MFMA kernels.
//
// Precision: the fp32 master parameters are rounded to bf16 when a block
// stages them in LDS; activations and back-propagated errors are rounded
// to bf16 where they enter a matrix product; every product accumulates in
// fp32 (v_mfma_f32_16x16x16_bf16); biases, ReLU masks and the loss stay
// fp32.  The optimiser (Adam) updates the fp32 masters.
//
// Work layout: a wavefront owns 16 rows (one MFMA tile) at a time.  With
// the 16x16x16 MFMA, lane l = r + 16 g (r = l % 16, g = l / 16) supplies
// A[r][4g..4g+3] and B[4g..4g+3][r] and receives C[4g..4g+3][r].  Forward
// activations go through LDS twice (row-major as the next A operand,
// transposed as the A operand of the weight gradient); the C-layout
// registers of an error tile are directly the B operand of the weight
// gradient, so dW += H^T dZ needs no extra shuffle.
//
// Training is deterministic: each wavefront keeps its partial gradient in
// registers over a fixed set of tiles, writes it to its own slot, and a
// second kernel sums the slots in slot order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kcl_internal.h"

namespace kcl {
namespace cm {

constexpr int F = KCL_CM_FEATURES;
constexpr int H = KCL_CM_HIDDEN;
constexpr int O = KCL_CM_OUTPUTS;
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + F * H;
constexpr int oW2 = oB1 + H;
constexpr int oB2 = oW2 + H * H;
constexpr int oW3 = oB2 + H;
constexpr int oB3 = oW3 + H * O;
constexpr int P = oB3 + O;
static_assert(P == KCL_CM_PARAMS, "parameter count");
static_assert(F <= 16 && O <= 16 && H == 64, "tile shapes");
constexpr int kWaves = 4;  // wavefronts per block
constexpr int kSlot = P + 1;  // partial gradient + loss sum

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (unsigned short)(u >> 16);
}

__device__ inline bf4 frag(const unsigned short *p) { return *(const bf4 *)p; }

__device__ inline bf4 pack(float a, float b, float c, float d) {
  bf4 r;
  r[0] = (short)to_bf16(a);
  r[1] = (short)to_bf16(b);
  r[2] = (short)to_bf16(c);
  r[3] = (short)to_bf16(d);
  return r;
}

__device__ inline f4 mfma(bf4 a, bf4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct WeightLds {
  unsigned short w1t[H][16];  // [i][f], f >= F zero
  unsigned short w2t[H][H];   // [j][i] = W2[i][j]
  unsigned short w2r[H][H];   // [i][j] = W2[i][j]
  unsigned short w3t[16][H];  // [c][j] = W3[j][c], c >= O zero
  unsigned short w3r[H][16];  // [j][c] = W3[j][c], c >= O zero
  float b1[H], b2[H], b3[16];
  float mu[16], inv_sd[16], ymu[16], ysd[16];
};

struct TileLds {
  unsigned short xt[16][16];  // [f][row]
  unsigned short h1[16][H];   // [row][i]
  unsigned short h1t[H][16];  // [i][row]
  unsigned short h2[16][H];
  unsigned short h2t[H][16];
  unsigned short dz2[16][H];  // [row][j]
  unsigned short dout[16][16];  // [row][c]
};

// rows: rows of this launch; row r reads input row idx ? idx[r] : r.
// Each wavefront handles tiles [w*tpw, (w+1)*tpw).  Predict (TRAIN false):
// out[r][c] de-normalised.  Train: slot w of `partial` receives the
// wavefront's gradient of sum_r,c (out - y)^2 / (2 inv_n^-1) and its
// squared-error sum.
template <bool TRAIN>
__global__ __launch_bounds__(64 * kWaves) void mlp_kernel(
    const float *__restrict__ params, const float *__restrict__ norm, const float *__restrict__ x,
    const float *__restrict__ y, const uint32_t *__restrict__ idx, uint32_t rows, uint32_t tpw,
    float inv_n, float *__restrict__ out, float *__restrict__ partial) {
  __shared__ WeightLds W;
  __shared__ TileLds T[kWaves];
  const int tid = threadIdx.x;
  for (int e = tid; e < H * 16; e += 64 * kWaves) {
    const int i = e / 16, f = e % 16;
    W.w1t[i][f] = f < F ? to_bf16(params[oW1 + f * H + i]) : 0;
    const int j = e / 16, c = e % 16;
    W.w3r[j][c] = c < O ? to_bf16(params[oW3 + j * O + c]) : 0;
    const int c2 = e / H, j2 = e % H;
    W.w3t[c2][j2] = c2 < O ? to_bf16(params[oW3 + j2 * O + c2]) : 0;
  }
  for (int e = tid; e < H * H; e += 64 * kWaves) {
    const int a = e / H, b = e % H;
    W.w2r[a][b] = to_bf16(params[oW2 + a * H + b]);
    W.w2t[a][b] = to_bf16(params[oW2 + b * H + a]);
  }
  if (tid < H) {
    W.b1[tid] = params[oB1 + tid];
    W.b2[tid] = params[oB2 + tid];
  }
  if (tid < 16) {
    W.b3[tid] = tid < O ? params[oB3 + tid] : 0.f;
    W.mu[tid] = tid < F ? norm[tid] : 0.f;
    W.inv_sd[tid] = tid < F ? 1.f / norm[F + tid] : 0.f;
    W.ymu[tid] = tid < O ? norm[2 * F + tid] : 0.f;
    W.ysd[tid] = tid < O ? norm[2 * F + O + tid] : 1.f;
  }
  __syncthreads();

  const int lane = tid & 63, wv = tid >> 6;
  const int r = lane & 15, g = lane >> 4;
  TileLds &L = T[wv];
  const uint64_t wave = (uint64_t)blockIdx.x * kWaves + wv;

  f4 gw1[4], gw2[4][4], gw3[4];
  float gb1[4], gb2[4], gb3 = 0.f, loss = 0.f;
  for (int a = 0; a < 4; ++a) {
    gw1[a] = f4{0, 0, 0, 0};
    gw3[a] = f4{0, 0, 0, 0};
    gb1[a] = gb2[a] = 0.f;
    for (int b = 0; b < 4; ++b) gw2[a][b] = f4{0, 0, 0, 0};
  }

  for (uint32_t tt = 0; tt < tpw; ++tt) {
    const uint64_t row0 = (wave * tpw + tt) * 16;
    __syncthreads();  // the previous tile's LDS reads are done
    // ---- inputs: lane (r, g) holds features 4g..4g+3 of row r ----
    const uint64_t my_row = row0 + r;
    const bool valid = my_row < rows;
    const uint64_t src = valid ? (idx ? idx[my_row] : my_row) : 0;
    float xv[4];
    for (int i = 0; i < 4; ++i) {
      const int f = 4 * g + i;
      xv[i] = (valid && f < F) ? (x[src * F + f] - W.mu[f]) * W.inv_sd[f] : 0.f;
      if (TRAIN) L.xt[f][r] = to_bf16(xv[i]);
    }
    const bf4 xa = pack(xv[0], xv[1], xv[2], xv[3]);
    // ---- layer 1: h1 = relu(x W1 + b1) ----
    f4 h1[4];
    for (int nb = 0; nb < 4; ++nb) {
      h1[nb] = mfma(xa, frag(&W.w1t[nb * 16 + r][4 * g]), f4{0, 0, 0, 0});
      const int col = nb * 16 + r;
      for (int i = 0; i < 4; ++i) {
        const float v = fmaxf(h1[nb][i] + W.b1[col], 0.f);
        h1[nb][i] = v;
        const unsigned short b = to_bf16(v);
        L.h1[4 * g + i][col] = b;
        L.h1t[col][4 * g + i] = b;
      }
    }
    __syncthreads();
    // ---- layer 2: h2 = relu(h1 W2 + b2) ----
    f4 h2[4];
    for (int nb = 0; nb < 4; ++nb) {
      f4 acc = f4{0, 0, 0, 0};
      for (int kb = 0; kb < 4; ++kb)
        acc = mfma(frag(&L.h1[r][kb * 16 + 4 * g]), frag(&W.w2t[nb * 16 + r][kb * 16 + 4 * g]),
                   acc);
      const int col = nb * 16 + r;
      for (int i = 0; i < 4; ++i) {
        const float v = fmaxf(acc[i] + W.b2[col], 0.f);
        h2[nb][i] = v;
        const unsigned short b = to_bf16(v);
        L.h2[4 * g + i][col] = b;
        L.h2t[col][4 * g + i] = b;
      }
    }
    __syncthreads();
    // ---- layer 3: out = h2 W3 + b3 (columns c >= O are padding) ----
    f4 o = f4{0, 0, 0, 0};
    for (int kb = 0; kb < 4; ++kb)
      o = mfma(frag(&L.h2[r][kb * 16 + 4 * g]), frag(&W.w3t[r][kb * 16 + 4 * g]), o);
    if (!TRAIN) {
      for (int i = 0; i < 4; ++i) {
        const uint64_t row = row0 + 4 * g + i;
        if (r < O && row < rows) out[row * O + r] = (o[i] + W.b3[r]) * W.ysd[r] + W.ymu[r];
      }
      continue;
    }
    // ---- loss and dL/dout = (out - y)/rows ----
    float d[4];
    for (int i = 0; i < 4; ++i) {
      const uint64_t row = row0 + 4 * g + i;
      float e = 0.f;
      if (r < O && row < rows) {
        const uint64_t sr = idx ? idx[row] : row;
        const float target = (y[sr * O + r] - W.ymu[r]) / W.ysd[r];
        e = (o[i] + W.b3[r]) - target;
      }
      loss += e * e;
      d[i] = e * inv_n;
      gb3 += d[i];
      L.dout[4 * g + i][r] = to_bf16(d[i]);
    }
    const bf4 db = pack(d[0], d[1], d[2], d[3]);
    for (int mb = 0; mb < 4; ++mb)  // dW3 += h2^T dout
      gw3[mb] = mfma(frag(&L.h2t[mb * 16 + r][4 * g]), db, gw3[mb]);
    __syncthreads();
    // ---- dz2 = (dout W3^T) * [h2 > 0];  dW2 += h1^T dz2 ----
    for (int nb = 0; nb < 4; ++nb) {
      const f4 dh = mfma(frag(&L.dout[r][4 * g]), frag(&W.w3r[nb * 16 + r][4 * g]),
                         f4{0, 0, 0, 0});
      float dz[4];
      const int col = nb * 16 + r;
      for (int i = 0; i < 4; ++i) {
        dz[i] = h2[nb][i] > 0.f ? dh[i] : 0.f;
        gb2[nb] += dz[i];
        L.dz2[4 * g + i][col] = to_bf16(dz[i]);
      }
      const bf4 dzb = pack(dz[0], dz[1], dz[2], dz[3]);
      for (int mb = 0; mb < 4; ++mb)
        gw2[mb][nb] = mfma(frag(&L.h1t[mb * 16 + r][4 * g]), dzb, gw2[mb][nb]);
    }
    __syncthreads();
    // ---- dz1 = (dz2 W2^T) * [h1 > 0];  dW1 += x^T dz1 ----
    for (int nb = 0; nb < 4; ++nb) {
      f4 acc = f4{0, 0, 0, 0};
      for (int kb = 0; kb < 4; ++kb)
        acc = mfma(frag(&L.dz2[r][kb * 16 + 4 * g]), frag(&W.w2r[nb * 16 + r][kb * 16 + 4 * g]),
                   acc);
      float dz[4];
      for (int i = 0; i < 4; ++i) {
        dz[i] = h1[nb][i] > 0.f ? acc[i] : 0.f;
        gb1[nb] += dz[i];
      }
      gw1[nb] = mfma(frag(&L.xt[r][4 * g]), pack(dz[0], dz[1], dz[2], dz[3]), gw1[nb]);
    }
  }
  if (!TRAIN) return;
  // ---- this wavefront's slot ----
  float *slot = partial + wave * kSlot;
  for (int nb = 0; nb < 4; ++nb)
    for (int i = 0; i < 4; ++i) {
      const int f = 4 * g + i;
      if (f < F) slot[oW1 + f * H + nb * 16 + r] = gw1[nb][i];
    }
  for (int mb = 0; mb < 4; ++mb)
    for (int nb = 0; nb < 4; ++nb)
      for (int i = 0; i < 4; ++i)
        slot[oW2 + (mb * 16 + 4 * g + i) * H + nb * 16 + r] = gw2[mb][nb][i];
  for (int mb = 0; mb < 4; ++mb)
    for (int i = 0; i < 4; ++i)
      if (r < O) slot[oW3 + (mb * 16 + 4 * g + i) * O + r] = gw3[mb][i];
  for (int nb = 0; nb < 4; ++nb) {
    float a = gb1[nb], b = gb2[nb];
    a += __shfl_xor(a, 16);
    a += __shfl_xor(a, 32);
    b += __shfl_xor(b, 16);
    b += __shfl_xor(b, 32);
    if (g == 0) {
      slot[oB1 + nb * 16 + r] = a;
      slot[oB2 + nb * 16 + r] = b;
    }
  }
  gb3 += __shfl_xor(gb3, 16);
  gb3 += __shfl_xor(gb3, 32);
  if (g == 0 && r < O) slot[oB3 + r] = gb3;
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (lane == 0) slot[P] = loss;
}

// grad[p] = sum over slots in slot order (deterministic).
__global__ void reduce_slots_kernel(const float *__restrict__ partial, uint32_t slots,
                                    float *__restrict__ grad) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kSlot) return;
  float s = 0.f;
  for (uint32_t w = 0; w < slots; ++w) s += partial[(size_t)w * kSlot + p];
  grad[p] = s;
}

// Adam step on the fp32 masters; step size lr_t carries the bias
// corrections.  epoch_loss += this batch's squared-error sum.
__global__ void adam_kernel(float *__restrict__ params, float *__restrict__ m1,
                            float *__restrict__ m2, const float *__restrict__ grad, float lr_t,
                            float b1, float b2, float eps, float *__restrict__ epoch_loss) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) epoch_loss[0] += grad[P];
  if (p >= (uint32_t)P) return;
  const float gr = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * gr;
  const float v = b2 * m2[p] + (1.f - b2) * gr * gr;
  m1[p] = a;
  m2[p] = v;
  params[p] -= lr_t * a / (sqrtf(v) + eps);
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace cm
}  // namespace kcl

struct kcl_cost_model {
  std::vector<float> params = std::vector<float>(kcl::cm::P, 0.f);
  std::vector<float> norm = std::vector<float>(KCL_CM_NORM, 0.f);
};

namespace kcl {
namespace cm {

void identity_norm(std::vector<float> &norm) {
  for (int i = 0; i < F; ++i) {
    norm[i] = 0.f;
    norm[F + i] = 1.f;
  }
  for (int c = 0; c < O; ++c) {
    norm[2 * F + c] = 0.f;
    norm[2 * F + O + c] = 1.f;
  }
}

kcl_status check_finite(const float *v, uint64_t n, const char *what) {
  for (uint64_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return fail(KCL_E_INVALID_ARG, "%s[%llu] is not finite", what,
                                          (unsigned long long)i);
  return KCL_OK;
}

// Device copy of a model for one call.
struct DevModel {
  DevBuf<float> params, norm;
  kcl_status upload(const kcl_cost_model *m, hipStream_t st) {
    KCL_TRY(params.alloc(P));
    KCL_TRY(norm.alloc(KCL_CM_NORM));
    KCL_HIP(hipMemcpyAsync(params.p, m->params.data(), P * sizeof(float), hipMemcpyHostToDevice,
                           st));
    KCL_HIP(hipMemcpyAsync(norm.p, m->norm.data(), KCL_CM_NORM * sizeof(float),
                           hipMemcpyHostToDevice, st));
    return KCL_OK;
  }
};

constexpr uint32_t kMaxBlocks = 512;

// Blocks and tiles per wavefront for `rows` rows.
void shape(uint64_t rows, uint32_t *blocks, uint32_t *tpw) {
  const uint64_t tiles = (rows + 15) / 16;
  uint64_t b = (tiles + kWaves - 1) / kWaves;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b == 0) b = 1;
  const uint64_t waves = b * kWaves;
  *blocks = (uint32_t)b;
  *tpw = (uint32_t)((tiles + waves - 1) / waves);
  if (*tpw == 0) *tpw = 1;
}

kcl_status predict_device(const kcl_cost_model *m, const float *d_x, uint64_t rows, float *d_y,
                          hipStream_t st) {
  if (rows == 0) return KCL_OK;
  if (rows > 0xffffffffull) return fail(KCL_E_INVALID_ARG, "too many rows");
  DevModel dm;
  KCL_TRY(dm.upload(m, st));
  uint32_t blocks, tpw;
  shape(rows, &blocks, &tpw);
  mlp_kernel<false><<<blocks, 64 * kWaves, 0, st>>>(dm.params.p, dm.norm.p, d_x, nullptr, nullptr,
                                                    (uint32_t)rows, tpw, 0.f, d_y, nullptr);
  KCL_HIP(hipGetLastError());
  KCL_HIP(hipStreamSynchronize(st));
  return KCL_OK;
}

}  // namespace cm

// Predictions for device-resident feature rows (used by the planner).
kcl_status cost_model_predict_device(const kcl_cost_model *m, const float *d_x, uint64_t rows,
                                     float *d_y, hipStream_t st) {
  return cm::predict_device(m, d_x, rows, d_y, st);
}

}  // namespace kcl

using namespace kcl;
using namespace kcl::cm;

extern "C" kcl_status kcl_cost_model_create(uint64_t seed, kcl_cost_model **out) {
  if (!out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  kcl_cost_model *m = new (std::nothrow) kcl_cost_model();
  if (!m) return fail(KCL_E_OOM_HOST, "out of host memory");
  uint64_t s = seed;
  auto uni = [&](float bound) {
    const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * bound);
  };
  for (int i = 0; i < F * H; ++i) m->params[oW1 + i] = uni(1.f / std::sqrt((float)F));
  for (int i = 0; i < H * H; ++i) m->params[oW2 + i] = uni(1.f / std::sqrt((float)H));
  for (int i = 0; i < H * O; ++i) m->params[oW3 + i] = uni(1.f / std::sqrt((float)H));
  identity_norm(m->norm);
  *out = m;
  return KCL_OK;
}

extern "C" void kcl_cost_model_destroy(kcl_cost_model *m) { delete m; }

static const char kMagic[8] = {'K', 'C', 'L', 'C', 'M', '0', '0', '1'};

extern "C" kcl_status kcl_cost_model_save(const kcl_cost_model *m, const char *path) {
  if (!m || !path) return fail(KCL_E_INVALID_ARG, "NULL argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(KCL_E_IO, "cannot open %s for writing", path);
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 &&
            std::fwrite(m->norm.data(), sizeof(float), KCL_CM_NORM, f) == KCL_CM_NORM &&
            std::fwrite(m->params.data(), sizeof(float), P, f) == (size_t)P;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? KCL_OK : fail(KCL_E_IO, "write to %s failed", path);
}

extern "C" kcl_status kcl_cost_model_load(const char *path, kcl_cost_model **out) {
  if (!path || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(KCL_E_IO, "cannot open %s", path);
  char magic[8];
  kcl_cost_model *m = new (std::nothrow) kcl_cost_model();
  if (!m) {
    std::fclose(f);
    return fail(KCL_E_OOM_HOST, "out of host memory");
  }
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                  std::fread(m->norm.data(), sizeof(float), KCL_CM_NORM, f) == KCL_CM_NORM &&
                  std::fread(m->params.data(), sizeof(float), P, f) == (size_t)P &&
                  std::fgetc(f) == EOF;
  std::fclose(f);
  kcl_status st = ok ? KCL_OK : fail(KCL_E_PARSE, "%s is not a cost model file", path);
  if (st == KCL_OK) st = check_finite(m->params.data(), P, "params");
  if (st == KCL_OK) st = check_finite(m->norm.data(), KCL_CM_NORM, "norm");
  if (st == KCL_OK)
    for (int i = 0; i < F; ++i)
      if (!(m->norm[F + i] > 0.f)) st = fail(KCL_E_PARSE, "%s: feature std %d <= 0", path, i);
  if (st != KCL_OK) {
    delete m;
    return st;
  }
  *out = m;
  return KCL_OK;
}

extern "C" kcl_status kcl_cost_model_get(const kcl_cost_model *m, float *params, float *norm) {
  if (!m) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (params) std::memcpy(params, m->params.data(), P * sizeof(float));
  if (norm) std::memcpy(norm, m->norm.data(), KCL_CM_NORM * sizeof(float));
  return KCL_OK;
}

extern "C" kcl_status kcl_cost_model_set(kcl_cost_model *m, const float *params, const float *norm) {
  if (!m) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (params) KCL_TRY(check_finite(params, P, "params"));
  if (norm) {
    KCL_TRY(check_finite(norm, KCL_CM_NORM, "norm"));
    for (int i = 0; i < F; ++i)
      if (!(norm[F + i] > 0.f)) return fail(KCL_E_INVALID_ARG, "feature std %d <= 0", i);
    for (int c = 0; c < O; ++c)
      if (!(norm[2 * F + O + c] > 0.f)) return fail(KCL_E_INVALID_ARG, "target std %d <= 0", c);
  }
  if (params) std::memcpy(m->params.data(), params, P * sizeof(float));
  if (norm) std::memcpy(m->norm.data(), norm, KCL_CM_NORM * sizeof(float));
  return KCL_OK;
}

namespace {

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace

extern "C" kcl_status kcl_cost_model_predict(kcl_cost_model *m, const float *x, uint64_t rows,
                                             float *y) {
  if (!m || (rows && (!x || !y))) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (rows == 0) return KCL_OK;
  Stream st;
  KCL_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  DevBuf<float> dx, dy;
  KCL_TRY(dx.alloc(rows * F));
  KCL_TRY(dy.alloc(rows * O));
  KCL_HIP(hipMemcpyAsync(dx.p, x, rows * F * sizeof(float), hipMemcpyHostToDevice, st.s));
  KCL_TRY(predict_device(m, dx.p, rows, dy.p, st.s));
  KCL_HIP(hipMemcpy(y, dy.p, rows * O * sizeof(float), hipMemcpyDeviceToHost));
  return KCL_OK;
}

extern "C" kcl_status kcl_cost_model_gradient(kcl_cost_model *m, const float *x, const float *y,
                                              uint64_t rows, float *loss, float *grad) {
  if (!m || !x || !y || rows == 0) return fail(KCL_E_INVALID_ARG, "NULL argument or no rows");
  if (rows > 0xffffffffull) return fail(KCL_E_INVALID_ARG, "too many rows");
  Stream st;
  KCL_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  DevModel dm;
  KCL_TRY(dm.upload(m, st.s));
  DevBuf<float> dx, dy, part, dg;
  KCL_TRY(dx.alloc(rows * F));
  KCL_TRY(dy.alloc(rows * O));
  KCL_HIP(hipMemcpyAsync(dx.p, x, rows * F * sizeof(float), hipMemcpyHostToDevice, st.s));
  KCL_HIP(hipMemcpyAsync(dy.p, y, rows * O * sizeof(float), hipMemcpyHostToDevice, st.s));
  uint32_t blocks, tpw;
  shape(rows, &blocks, &tpw);
  const uint32_t slots = blocks * kWaves;
  KCL_TRY(part.alloc((size_t)slots * kSlot));
  KCL_TRY(dg.alloc(kSlot));
  mlp_kernel<true><<<blocks, 64 * kWaves, 0, st.s>>>(dm.params.p, dm.norm.p, dx.p, dy.p, nullptr,
                                                     (uint32_t)rows, tpw, 1.f / (float)rows,
                                                     nullptr, part.p);
  KCL_HIP(hipGetLastError());
  reduce_slots_kernel<<<(kSlot + 255) / 256, 256, 0, st.s>>>(part.p, slots, dg.p);
  KCL_HIP(hipGetLastError());
  std::vector<float> h(kSlot);
  KCL_HIP(hipMemcpyAsync(h.data(), dg.p, kSlot * sizeof(float), hipMemcpyDeviceToHost, st.s));
  KCL_HIP(hipStreamSynchronize(st.s));
  if (grad) std::memcpy(grad, h.data(), P * sizeof(float));
  if (loss) *loss = h[P] / (2.f * (float)rows);
  return KCL_OK;
}

extern "C" kcl_status kcl_cost_model_train(kcl_cost_model *m, const float *x, const float *y,
                                           uint64_t rows, const kcl_cm_train_opts *opts,
                                           float *loss_per_epoch) {
  if (!m || !x || !y || rows == 0) return fail(KCL_E_INVALID_ARG, "NULL argument or no rows");
  if (rows > 0xffffffffull) return fail(KCL_E_INVALID_ARG, "too many rows");
  kcl_cm_train_opts o{};
  if (opts) o = *opts;
  const uint32_t epochs = o.epochs ? o.epochs : 60;
  const uint32_t batch = o.batch ? o.batch : 2048;
  const float lr = o.lr > 0.f ? o.lr : 2e-3f;
  KCL_TRY(check_finite(x, rows * F, "x"));
  KCL_TRY(check_finite(y, rows * O, "y"));
  if (o.fit_norm) {
    std::vector<float> nn(KCL_CM_NORM);
    for (int f = 0; f < F + O; ++f) {
      const float *col = f < F ? x + f : y + (f - F);
      const int stride = f < F ? F : O;
      double s = 0, s2 = 0;
      for (uint64_t i = 0; i < rows; ++i) s += col[i * stride];
      const double mean = s / (double)rows;
      for (uint64_t i = 0; i < rows; ++i) {
        const double d = col[i * stride] - mean;
        s2 += d * d;
      }
      double sd = std::sqrt(s2 / (double)rows);
      if (!(sd > 1e-6)) sd = 1.0;  // constant column
      if (f < F) {
        nn[f] = (float)mean;
        nn[F + f] = (float)sd;
      } else {
        nn[2 * F + (f - F)] = (float)mean;
        nn[2 * F + O + (f - F)] = (float)sd;
      }
    }
    m->norm = nn;
  }
  Stream st;
  KCL_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  DevModel dm;
  KCL_TRY(dm.upload(m, st.s));
  DevBuf<float> dx, dy, part, dg, m1, m2, eloss;
  DevBuf<uint32_t> didx;
  KCL_TRY(dx.alloc(rows * F));
  KCL_TRY(dy.alloc(rows * O));
  KCL_TRY(didx.alloc(rows));
  KCL_TRY(m1.alloc(P));
  KCL_TRY(m2.alloc(P));
  KCL_TRY(dg.alloc(kSlot));
  KCL_TRY(eloss.alloc(epochs));
  KCL_HIP(hipMemcpyAsync(dx.p, x, rows * F * sizeof(float), hipMemcpyHostToDevice, st.s));
  KCL_HIP(hipMemcpyAsync(dy.p, y, rows * O * sizeof(float), hipMemcpyHostToDevice, st.s));
  KCL_HIP(hipMemsetAsync(m1.p, 0, P * sizeof(float), st.s));
  KCL_HIP(hipMemsetAsync(m2.p, 0, P * sizeof(float), st.s));
  KCL_HIP(hipMemsetAsync(eloss.p, 0, epochs * sizeof(float), st.s));
  uint32_t max_blocks, max_tpw;
  shape(std::min<uint64_t>(rows, batch), &max_blocks, &max_tpw);
  KCL_TRY(part.alloc((size_t)max_blocks * kWaves * kSlot));
  std::vector<uint32_t> perm(rows);
  uint64_t rng = o.seed ^ 0x6b636c636d747261ull;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  uint64_t step = 0;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = 0; i < rows; ++i) perm[i] = (uint32_t)i;
    for (uint64_t i = rows - 1; i > 0; --i) {  // Fisher-Yates
      const uint64_t j = splitmix64(rng) % (i + 1);
      std::swap(perm[i], perm[j]);
    }
    // The previous epoch's launches read perm's device copy: wait for them.
    KCL_HIP(hipStreamSynchronize(st.s));
    KCL_HIP(hipMemcpyAsync(didx.p, perm.data(), rows * sizeof(uint32_t), hipMemcpyHostToDevice,
                           st.s));
    for (uint64_t b0 = 0; b0 < rows; b0 += batch) {
      const uint64_t cnt = std::min<uint64_t>(batch, rows - b0);
      uint32_t blocks, tpw;
      shape(cnt, &blocks, &tpw);
      mlp_kernel<true><<<blocks, 64 * kWaves, 0, st.s>>>(dm.params.p, dm.norm.p, dx.p, dy.p,
                                                         didx.p + b0, (uint32_t)cnt, tpw,
                                                         1.f / (float)cnt, nullptr, part.p);
      KCL_HIP(hipGetLastError());
      reduce_slots_kernel<<<(kSlot + 255) / 256, 256, 0, st.s>>>(part.p, blocks * kWaves, dg.p);
      KCL_HIP(hipGetLastError());
      ++step;
      const double c1 = 1.0 - std::pow((double)b1, (double)step);
      const double c2 = 1.0 - std::pow((double)b2, (double)step);
      const float lr_t = (float)(lr * std::sqrt(c2) / c1);
      adam_kernel<<<(P + 255) / 256, 256, 0, st.s>>>(dm.params.p, m1.p, m2.p, dg.p, lr_t, b1, b2,
                                                     eps, eloss.p + ep);
      KCL_HIP(hipGetLastError());
    }
  }
  std::vector<float> el(epochs);
  KCL_HIP(hipMemcpyAsync(m->params.data(), dm.params.p, P * sizeof(float), hipMemcpyDeviceToHost,
                         st.s));
  KCL_HIP(hipMemcpyAsync(el.data(), eloss.p, epochs * sizeof(float), hipMemcpyDeviceToHost, st.s));
  KCL_HIP(hipStreamSynchronize(st.s));
  if (loss_per_epoch)
    for (uint32_t e = 0; e < epochs; ++e) loss_per_epoch[e] = el[e] / (2.f * (float)rows);
  return check_finite(m->params.data(), P, "trained params");
}

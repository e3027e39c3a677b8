// xf_lb_train.inl -- training the learned absorbing layer's column network
// on the GPU (no reference counterpart).  Included at the end of
// xf_engine.hip.
//
// Data: raw samples (12 inputs + the next layer-0 value, recorded by the
// engine in XF_LB_RECORD mode) are turned on the device into the network's
// scaled inputs x^ = x / s and residual targets r^ = (y - mur) / s, with s
// and mur exactly as the engine's predictor forms them.
//
// Gradient: one thread per sample runs the forward pass with the fp16-
// rounded weights (fp32 arithmetic) and the backward pass of the mean
// squared error; every one of the 1505 parameter gradients (and the loss) is
// summed over the wave's 64 samples with an xor butterfly, accumulated per
// wave in LDS, summed over the block's waves and written as a per-block
// partial; a second kernel adds the partials in block order.  Every sum has
// a fixed order, so a gradient is reproducible bit for bit.
//
// Optimiser: Adam on fp32 master weights; the forward pass sees them
// rounded to fp16 (the layer's inference precision), the update is applied
// to the fp32 copy (straight-through).
#include <hip/hip_fp16.h>

namespace xf {

constexpr int kLbWaves = 4;  // waves per block of the gradient kernel
constexpr int kLbAcc = XF_LB_NPARAM + 1;  // gradient + loss

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// The network's inputs and target of one raw sample (13 floats): x^ = x / s,
// r^ = (y - mur) / s with s = 2^e the power of two above the largest |x|
// (frexp), exactly as the predictor scales them; e = INT_MIN / 4 marks an
// all-zero sample.
__device__ __forceinline__ void lb_prep_one(const float* __restrict__ x, float mur_k, float* xs, float& r,
                                            int& ex) {
  float m = 0.0f;
  for (int i = 0; i < XF_LB_IN; ++i) {
    const float a = fabsf(x[i]);
    if (a > m) m = a;
  }
  const float mur = x[4] + mur_k * (x[0] - x[3]);
  int e = 0;
  if (m > 0.0f) (void)frexpf(m, &e);
  for (int i = 0; i < XF_LB_IN; ++i) xs[i] = m > 0.0f ? ldexpf(x[i], -e) : 0.0f;
  r = m > 0.0f ? ldexpf(x[12] - mur, -e) : 0.0f;
  ex = m > 0.0f ? e : INT_MIN / 4;
}

// Loss weight (s / s_max)^power of a sample whose input scale is s = 2^ex
// (0 for an all-zero sample; the exponent difference is formed in 64 bits).
__device__ __forceinline__ float lb_weight(int ex, int e_max, int power) {
  const int64_t d = (int64_t)power * ((int64_t)ex - e_max);
  return d < -126 ? 0.0f : ldexpf(1.0f, (int)d);
}

__global__ void __launch_bounds__(256)
    k_lb_prep(const float* __restrict__ raw, int64_t n, float mur_k, float* __restrict__ xs,
              float* __restrict__ r, int* __restrict__ ex) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float x[XF_LB_IN];
  float rk;
  int e;
  lb_prep_one(raw + k * 13, mur_k, x, rk, e);
  for (int i = 0; i < XF_LB_IN; ++i) xs[k * XF_LB_IN + i] = x[i];
  r[k] = rk;
  ex[k] = e;
}

// Loss weight of each sample.
__global__ void k_lb_weight(const int* __restrict__ ex, int64_t n, int e_max, int power,
                            float* __restrict__ lw) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  lw[k] = lb_weight(ex[k], e_max, power);
}

// Counter-based minibatch indices (splitmix64 of seed, step, k).
__global__ void k_lb_sample(uint32_t* __restrict__ idx, int64_t count, int64_t n, uint64_t seed,
                            int64_t step) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(step * 1000003 + k + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  idx[k] = (uint32_t)(z % (uint64_t)n);
}

// Where the gradient kernels find a sample's scaled inputs, target and loss
// weight: the trainer's prepared arrays ...
struct LbPrepared {
  const float* __restrict__ xs;
  const float* __restrict__ r;
  const float* __restrict__ lw;
  __device__ __forceinline__ void get(int64_t s, float* x, float& target, float& w) const {
#pragma unroll
    for (int i = 0; i < XF_LB_IN; ++i) x[i] = xs[s * XF_LB_IN + i];
    target = r[s];
    w = lw[s];
  }
};

// ... or the raw samples, prepared on the fly by the same arithmetic
// (lb_prep_one, lb_weight), with the largest |input| over all samples as
// float bits at *absmax_bits.
struct LbRaw {
  const float* __restrict__ raw;
  float mur_k;
  int power;
  const unsigned* __restrict__ absmax_bits;
  __device__ __forceinline__ void get(int64_t s, float* x, float& target, float& w) const {
    int e;
    lb_prep_one(raw + s * 13, mur_k, x, target, e);
    const float smax = __uint_as_float(*absmax_bits);
    int e_max = 0;
    if (smax > 0.0f) (void)frexpf(smax, &e_max);
    w = lb_weight(e, e_max, power);
  }
};

// Loss and gradient over `count` samples: sample k is idx[k] (or first + k
// without an index list); the loss is sum d^2 * inv_count.  The weights are
// the fp16 rounding of w (RoundW: w is fp32 master weights; otherwise w
// already holds the rounded values).
template <typename Src, bool RoundW>
__device__ __forceinline__ void lb_grad_body(const Src& src, const uint32_t* __restrict__ idx, int64_t first,
                                             int64_t count, float inv_count, const float* __restrict__ w_in,
                                             float* __restrict__ partial) {
  __shared__ float w[XF_LB_NPARAM];
  __shared__ float acc[kLbWaves][kLbAcc];
  for (int i = threadIdx.x; i < XF_LB_NPARAM; i += blockDim.x)
    w[i] = RoundW ? __half2float(__float2half_rn(w_in[i])) : w_in[i];
  for (int i = threadIdx.x; i < kLbWaves * kLbAcc; i += blockDim.x) (&acc[0][0])[i] = 0.0f;
  __syncthreads();
  const float* W1 = w;
  const float* b1 = W1 + XF_LB_HID * XF_LB_IN;
  const float* W2 = b1 + XF_LB_HID;
  const float* b2 = W2 + XF_LB_HID * XF_LB_HID;
  const float* w3 = b2 + XF_LB_HID;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  float* a = acc[wave];
  const int64_t stride = (int64_t)gridDim.x * kLbWaves * 64;
  for (int64_t base = ((int64_t)blockIdx.x * kLbWaves + wave) * 64; base < count; base += stride) {
    const int64_t k = base + lane;
    const bool valid = k < count;
    float x[XF_LB_IN];
    float target = 0.0f, lw = 0.0f;
    if (valid) {
      src.get(idx ? (int64_t)idx[k] : first + k, x, target, lw);
    } else {
#pragma unroll
      for (int i = 0; i < XF_LB_IN; ++i) x[i] = 0.0f;
    }
    float h1[XF_LB_HID], h2[XF_LB_HID];
#pragma unroll
    for (int o = 0; o < XF_LB_HID; ++o) {
      float z = b1[o];
#pragma unroll
      for (int i = 0; i < XF_LB_IN; ++i) z = z + W1[o * XF_LB_IN + i] * x[i];
      h1[o] = z > 0.0f ? z : 0.0f;
    }
    float out = w3[XF_LB_HID];
#pragma unroll
    for (int o = 0; o < XF_LB_HID; ++o) {
      float z = b2[o];
#pragma unroll
      for (int i = 0; i < XF_LB_HID; ++i) z = z + W2[o * XF_LB_HID + i] * h1[i];
      h2[o] = z > 0.0f ? z : 0.0f;
      out = out + w3[o] * h2[o];
    }
    const float d = out - target;
    const float g_out = valid ? 2.0f * lw * d * inv_count : 0.0f;
    const float l = valid ? lw * d * d * inv_count : 0.0f;
    // output layer
    float s;
#pragma unroll
    for (int o = 0; o < XF_LB_HID; ++o) {
      s = wave_sum(g_out * h2[o]);
      if (lane == 0) a[XF_LB_HID * XF_LB_IN + XF_LB_HID + XF_LB_HID * XF_LB_HID + XF_LB_HID + o] += s;
    }
    s = wave_sum(g_out);
    if (lane == 0) a[XF_LB_NPARAM - 1] += s;
    s = wave_sum(l);
    if (lane == 0) a[XF_LB_NPARAM] += s;
    // second hidden layer
    float gz2[XF_LB_HID];
#pragma unroll
    for (int o = 0; o < XF_LB_HID; ++o) gz2[o] = h2[o] > 0.0f ? g_out * w3[o] : 0.0f;
    const int oW2 = XF_LB_HID * XF_LB_IN + XF_LB_HID;
    const int ob2 = oW2 + XF_LB_HID * XF_LB_HID;
#pragma unroll
    for (int o = 0; o < XF_LB_HID; ++o) {
#pragma unroll
      for (int i = 0; i < XF_LB_HID; ++i) {
        s = wave_sum(gz2[o] * h1[i]);
        if (lane == 0) a[oW2 + o * XF_LB_HID + i] += s;
      }
      s = wave_sum(gz2[o]);
      if (lane == 0) a[ob2 + o] += s;
    }
    // first hidden layer
    const int ob1 = XF_LB_HID * XF_LB_IN;
#pragma unroll
    for (int i = 0; i < XF_LB_HID; ++i) {
      float g = 0.0f;
#pragma unroll
      for (int o = 0; o < XF_LB_HID; ++o) g = g + gz2[o] * W2[o * XF_LB_HID + i];
      const float gz1 = h1[i] > 0.0f ? g : 0.0f;
#pragma unroll
      for (int j = 0; j < XF_LB_IN; ++j) {
        s = wave_sum(gz1 * x[j]);
        if (lane == 0) a[i * XF_LB_IN + j] += s;
      }
      s = wave_sum(gz1);
      if (lane == 0) a[ob1 + i] += s;
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kLbAcc; p += blockDim.x) {
    float t = acc[0][p];
#pragma unroll
    for (int q = 1; q < kLbWaves; ++q) t += acc[q][p];
    partial[(int64_t)blockIdx.x * kLbAcc + p] = t;
  }
}

__global__ void __launch_bounds__(256)
    k_lb_grad(const float* __restrict__ xs_all, const float* __restrict__ r_all,
              const float* __restrict__ lw_all, const uint32_t* __restrict__ idx, int64_t first,
              int64_t count, float inv_count, const float* __restrict__ w16,
              float* __restrict__ partial) {
  lb_grad_body<LbPrepared, false>(LbPrepared{xs_all, r_all, lw_all}, idx, first, count, inv_count, w16,
                                  partial);
}

__global__ void __launch_bounds__(256)
    k_lb_grad_raw(const float* __restrict__ raw, float mur_k, int power,
                  const unsigned* __restrict__ absmax_bits, int64_t first, int64_t count, float inv_count,
                  const float* __restrict__ w32, float* __restrict__ partial) {
  lb_grad_body<LbRaw, true>(LbRaw{raw, mur_k, power, absmax_bits}, nullptr, first, count, inv_count, w32,
                            partial);
}

// Largest |input| over the 12 inputs of all n raw samples, as float bits
// (non-negative floats order like their bit patterns), folded into *bits,
// which the caller cleared.
__global__ void __launch_bounds__(256)
    k_lb_absmax(const float* __restrict__ raw, int64_t n, unsigned* __restrict__ bits) {
  __shared__ unsigned wmax[4];
  unsigned m = 0u;
  const int64_t total = n * 13;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x)
    if (q % 13 != 12) m = max(m, __float_as_uint(fabsf(raw[q])));
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, s, 64));
  if (threadIdx.x % 64 == 0) wmax[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = wmax[0];
    for (int w = 1; w < (int)(blockDim.x / 64); ++w) b = max(b, wmax[w]);
    atomicMax(bits, b);
  }
}

__global__ void k_lb_clear_u32(unsigned* __restrict__ p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

__global__ void k_lb_reduce(const float* __restrict__ partial, int nblocks, float* __restrict__ grad,
                            double* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kLbAcc) return;
  double t = 0.0;
  for (int b = 0; b < nblocks; ++b) t += (double)partial[(int64_t)b * kLbAcc + p];
  if (p < XF_LB_NPARAM)
    grad[p] = (float)t;
  else
    *loss = t;
}

__global__ void k_lb_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ g, float* __restrict__ w16, int64_t step, float lr,
                          float b1, float b2, float eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= XF_LB_NPARAM) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  const float c1 = 1.0f - powf(b1, (float)step);
  const float c2 = 1.0f - powf(b2, (float)step);
  const float wp = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
  w[p] = wp;
  w16[p] = __half2float(__float2half_rn(wp));
}

__global__ void k_lb_round(const float* __restrict__ w, float* __restrict__ w16,
                           uint16_t* __restrict__ bits) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= XF_LB_NPARAM) return;
  const __half h = __float2half_rn(w[p]);
  w16[p] = __half2float(h);
  if (bits) bits[p] = __half_as_ushort(h);
}

// Blocks of the gradient kernel on a device with `cus` compute units (the
// trainer and the no-allocation entry point use the same count, so their
// per-block partials, and hence their sums, are the same).
constexpr int kLbMaxBlocks = 1024;
inline int lb_grad_blocks(int cus) { return std::min(2 * cus, kLbMaxBlocks); }

struct LbTrainer {
  int device{0};
  hipStream_t stream{};
  std::vector<void*> allocs;
  int64_t n{0};
  float mur_k{0};
  float* xs{};
  float* r{};
  float* lw{};   // per-sample loss weight
  int* ex{};     // per-sample input-scale exponent
  int e_max{0};
  float* w{};     // fp32 master
  float* w16{};   // fp16-rounded copy used by the forward pass
  uint16_t* bits{};
  float* m{};
  float* v{};
  float* g{};
  double* loss{};
  float* partial{};
  uint32_t* idx{};
  int64_t idx_cap{0};
  int nblocks{0};
  int64_t step{0};

  ~LbTrainer() {
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int alloc(void** p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(XF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    return XF_OK;
  }

  int create(const float* samples, int64_t count, double k, int dev) {
    device = dev;
    n = count;
    mur_k = static_cast<float>(k);
    XF_HIP(hipSetDevice(device));
    XF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    void* p;
    XF_TRY(alloc(&p, static_cast<size_t>(n) * 13 * sizeof(float)));
    float* raw = static_cast<float*>(p);
    XF_HIP(hipMemcpy(raw, samples, static_cast<size_t>(n) * 13 * sizeof(float), hipMemcpyHostToDevice));
    XF_TRY(alloc(&p, static_cast<size_t>(n) * XF_LB_IN * sizeof(float)));
    xs = static_cast<float*>(p);
    XF_TRY(alloc(&p, static_cast<size_t>(n) * sizeof(float)));
    r = static_cast<float*>(p);
    XF_TRY(alloc(&p, static_cast<size_t>(n) * sizeof(float)));
    lw = static_cast<float*>(p);
    XF_TRY(alloc(&p, static_cast<size_t>(n) * sizeof(int)));
    ex = static_cast<int*>(p);
    hipLaunchKernelGGL(k_lb_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, raw, n, mur_k, xs, r,
                       ex);
    XF_HIP(hipGetLastError());
    float smax = 0.0f;
    for (int64_t q = 0; q < n * 13; ++q)
      if (q % 13 != 12) smax = std::max(smax, std::fabs(samples[q]));
    if (smax > 0.0f) (void)std::frexp(smax, &e_max);
    float** vecs[5] = {&w, &w16, &m, &v, &g};
    for (float** q : vecs) {
      XF_TRY(alloc(&p, XF_LB_NPARAM * sizeof(float)));
      *q = static_cast<float*>(p);
      XF_HIP(hipMemsetAsync(*q, 0, XF_LB_NPARAM * sizeof(float), stream));
    }
    XF_TRY(alloc(&p, XF_LB_NPARAM * sizeof(uint16_t)));
    bits = static_cast<uint16_t*>(p);
    XF_TRY(alloc(&p, sizeof(double)));
    loss = static_cast<double*>(p);
    hipDeviceProp_t prop;
    XF_HIP(hipGetDeviceProperties(&prop, device));
    nblocks = lb_grad_blocks(prop.multiProcessorCount);
    XF_TRY(alloc(&p, static_cast<size_t>(nblocks) * kLbAcc * sizeof(float)));
    partial = static_cast<float*>(p);
    XF_TRY(set_weighting(0));
    return XF_OK;
  }

  int set_weighting(int power) {
    if (power < 0 || power > 4) return fail(XF_ERR_INVALID, "weighting power must be 0..4");
    hipLaunchKernelGGL(k_lb_weight, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ex, n, e_max, power,
                       lw);
    XF_HIP(hipGetLastError());
    XF_HIP(hipStreamSynchronize(stream));
    return XF_OK;
  }

  int set_weights(const float* host) {
    XF_HIP(hipMemcpy(w, host, XF_LB_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_lb_round, dim3((XF_LB_NPARAM + 255) / 256), dim3(256), 0, stream, w, w16,
                       static_cast<uint16_t*>(nullptr));
    XF_HIP(hipMemsetAsync(m, 0, XF_LB_NPARAM * sizeof(float), stream));
    XF_HIP(hipMemsetAsync(v, 0, XF_LB_NPARAM * sizeof(float), stream));
    step = 0;
    XF_HIP(hipStreamSynchronize(stream));
    return XF_OK;
  }

  int get_weights(float* w32_host, uint16_t* bits_host) {
    hipLaunchKernelGGL(k_lb_round, dim3((XF_LB_NPARAM + 255) / 256), dim3(256), 0, stream, w, w16, bits);
    XF_HIP(hipStreamSynchronize(stream));
    if (w32_host) XF_HIP(hipMemcpy(w32_host, w, XF_LB_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    if (bits_host) XF_HIP(hipMemcpy(bits_host, bits, XF_LB_NPARAM * sizeof(uint16_t), hipMemcpyDeviceToHost));
    return XF_OK;
  }

  // loss and gradient over idx[0..count) (or samples [first, first+count))
  int launch_grad(const uint32_t* ix, int64_t first, int64_t count) {
    const int blocks = static_cast<int>(std::min<int64_t>(nblocks, (count + kLbWaves * 64 - 1) / (kLbWaves * 64)));
    hipLaunchKernelGGL(k_lb_grad, dim3(blocks), dim3(64 * kLbWaves), 0, stream, xs, r, lw, ix, first, count,
                       1.0f / static_cast<float>(count), w16, partial);
    hipLaunchKernelGGL(k_lb_reduce, dim3((kLbAcc + 255) / 256), dim3(256), 0, stream, partial, blocks, g, loss);
    XF_HIP(hipGetLastError());
    return XF_OK;
  }

  int grad(int64_t first, int64_t count, double* loss_host, float* grad_host) {
    if (first < 0 || count < 1 || first + count > n) return fail(XF_ERR_INVALID, "sample range out of bounds");
    XF_TRY(launch_grad(nullptr, first, count));
    XF_HIP(hipStreamSynchronize(stream));
    if (loss_host) XF_HIP(hipMemcpy(loss_host, loss, sizeof(double), hipMemcpyDeviceToHost));
    if (grad_host) XF_HIP(hipMemcpy(grad_host, g, XF_LB_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return XF_OK;
  }

  int train(int steps, int batch, float lr, uint64_t seed, double* last_loss) {
    if (steps < 0 || batch < 1) return fail(XF_ERR_INVALID, "bad training schedule");
    if (batch > idx_cap) {
      void* p;
      XF_TRY(alloc(&p, static_cast<size_t>(batch) * sizeof(uint32_t)));
      idx = static_cast<uint32_t*>(p);
      idx_cap = batch;
    }
    for (int s = 0; s < steps; ++s) {
      ++step;
      hipLaunchKernelGGL(k_lb_sample, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, stream, idx,
                         static_cast<int64_t>(batch), n, seed, step);
      XF_TRY(launch_grad(idx, 0, batch));
      hipLaunchKernelGGL(k_lb_adam, dim3((XF_LB_NPARAM + 255) / 256), dim3(256), 0, stream, w, m, v, g, w16,
                         step, lr, 0.9f, 0.999f, 1e-8f);
      XF_HIP(hipGetLastError());
    }
    XF_HIP(hipStreamSynchronize(stream));
    if (last_loss) XF_HIP(hipMemcpy(last_loss, loss, sizeof(double), hipMemcpyDeviceToHost));
    return XF_OK;
  }
};

// ---- The no-allocation gradient entry point.  Its working memory is this
// file-scope array, one copy per device: word 0 holds the largest |input|
// (float bits); the per-block partials start at float 64.  Its device
// address is looked up once per device with hipGetSymbolAddress and kept;
// host code never passes the array's own name to a kernel.
constexpr int64_t kLbScratchHead = 64;
constexpr int64_t kLbScratchFloats = kLbScratchHead + (int64_t)kLbMaxBlocks * kLbAcc;
static __device__ __align__(256) float scratch[kLbScratchFloats];

constexpr int kLbMaxDevices = 64;
struct LbScratchSlot {
  std::once_flag once;
  int status{XF_ERR_HIP};
  std::string err;
  float* ptr{nullptr};
  int cus{0};
};
LbScratchSlot g_lb_scratch[kLbMaxDevices];

// Set-up of one device's scratch (with that device current): its address
// and the device's compute-unit count.
int lb_scratch_setup(int device, float** ptr, int* cus) {
  LbScratchSlot& S = g_lb_scratch[device];
  std::call_once(S.once, [&]() {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      S.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    int n = 0;
    e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
      S.err = std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e);
      return;
    }
    S.ptr = static_cast<float*>(p);
    S.cus = n;
    S.status = XF_OK;
  });
  if (S.status != XF_OK) return fail(S.status, S.err);
  *ptr = S.ptr;
  *cus = S.cus;
  return XF_OK;
}

int lb_grad_noalloc(const float* raw, int64_t n, double mur_k, int power, const float* w32, int64_t first,
                    int64_t count, double* loss, float* grad, int device, hipStream_t stream) {
  if (!raw || !w32 || !loss || !grad) return fail(XF_ERR_INVALID, "null argument");
  if (n < 1) return fail(XF_ERR_INVALID, "no samples");
  if (n > (int64_t)UINT32_MAX) return fail(XF_ERR_UNSUPPORTED, "too many samples (32-bit indices)");
  if (first < 0 || count < 1 || first + count > n) return fail(XF_ERR_INVALID, "sample range out of bounds");
  if (power < 0 || power > 4) return fail(XF_ERR_INVALID, "weighting power must be 0..4");
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return fail(XF_ERR_HIP, "no HIP device");
  if (device < 0 || device >= nd) return fail(XF_ERR_INVALID, "bad device ordinal");
  if (device >= kLbMaxDevices) return fail(XF_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  XF_HIP(hipSetDevice(device));
  float* base = nullptr;
  int cus = 0;
  XF_TRY(lb_scratch_setup(device, &base, &cus));
  const int nblocks = lb_grad_blocks(cus);
  const int blocks = static_cast<int>(std::min<int64_t>(nblocks, (count + kLbWaves * 64 - 1) / (kLbWaves * 64)));
  if ((int64_t)kLbScratchHead + (int64_t)blocks * kLbAcc > kLbScratchFloats)
    return fail(XF_ERR_UNSUPPORTED, "gradient partials exceed the static scratch");
  unsigned* absmax = reinterpret_cast<unsigned*>(base);
  float* partial = base + kLbScratchHead;
  hipLaunchKernelGGL(k_lb_clear_u32, dim3(1), dim3(64), 0, stream, absmax);
  const int64_t total = n * 13;
  const unsigned mblocks = static_cast<unsigned>(std::min<int64_t>(4 * cus, (total + 255) / 256));
  hipLaunchKernelGGL(k_lb_absmax, dim3(mblocks), dim3(256), 0, stream, raw, n, absmax);
  hipLaunchKernelGGL(k_lb_grad_raw, dim3(blocks), dim3(64 * kLbWaves), 0, stream, raw,
                     static_cast<float>(mur_k), power, static_cast<const unsigned*>(absmax), first, count,
                     1.0f / static_cast<float>(count), w32, partial);
  hipLaunchKernelGGL(k_lb_reduce, dim3((kLbAcc + 255) / 256), dim3(256), 0, stream,
                     static_cast<const float*>(partial), blocks, grad, loss);
  XF_HIP(hipGetLastError());
  return XF_OK;
}

}  // namespace xf

struct xf_lb_trainer {
  xf::LbTrainer t;
};

extern "C" {

int xf_lb_trainer_create(const float* samples, int64_t n, double mur_k, int32_t device,
                         xf_lb_trainer** out) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!samples || !out || n < 1) return xf::fail(XF_ERR_INVALID, "no samples");
    if (n > (int64_t)UINT32_MAX) return xf::fail(XF_ERR_UNSUPPORTED, "too many samples (32-bit indices)");
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return xf::fail(XF_ERR_HIP, "no HIP device");
    if (device < 0 || device >= nd) return xf::fail(XF_ERR_INVALID, "bad device ordinal");
    std::unique_ptr<xf_lb_trainer> h(new (std::nothrow) xf_lb_trainer);
    if (!h) return xf::fail(XF_ERR_NOMEM, "out of host memory");
    const int st = h->t.create(samples, n, mur_k, device);
    if (st != XF_OK) return st;
    *out = h.release();
    return XF_OK;
  });
}

int xf_lb_trainer_set_weights(xf_lb_trainer* h, const float* w) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h || !w) return xf::fail(XF_ERR_INVALID, "null argument");
    return h->t.set_weights(w);
  });
}

int xf_lb_trainer_get_weights(xf_lb_trainer* h, float* w32, uint16_t* w16) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null trainer");
    return h->t.get_weights(w32, w16);
  });
}

int xf_lb_trainer_grad(xf_lb_trainer* h, int64_t first, int64_t count, double* loss, float* grad) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null trainer");
    return h->t.grad(first, count, loss, grad);
  });
}

int xf_lb_trainer_train(xf_lb_trainer* h, int32_t steps, int32_t batch, float lr, uint64_t seed,
                        double* last_loss) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null trainer");
    return h->t.train(steps, batch, lr, seed, last_loss);
  });
}

int xf_lb_trainer_samples(const xf_lb_trainer* h, int64_t* n) {
  return xf::guarded([&]() -> int {
    if (!h || !n) return xf::fail(XF_ERR_INVALID, "null argument");
    *n = h->t.n;
    return XF_OK;
  });
}

int xf_lb_trainer_set_weighting(xf_lb_trainer* h, int32_t power) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    if (!h) return xf::fail(XF_ERR_INVALID, "null trainer");
    return h->t.set_weighting(power);
  });
}

int xf_lb_grad_noalloc(const float* raw, int64_t n, double mur_k, int32_t power, const float* w32,
                       int64_t first, int64_t count, double* loss, float* grad, int32_t device,
                       void* stream) {
  return xf::guarded([&]() -> int {
    xf::g_err.clear();
    return xf::lb_grad_noalloc(raw, n, mur_k, power, w32, first, count, loss, grad, device,
                               static_cast<hipStream_t>(stream));
  });
}

int xf_lb_trainer_destroy(xf_lb_trainer* h) {
  return xf::guarded([&]() -> int {
    delete h;
    return XF_OK;
  });
}

}  // extern "C"

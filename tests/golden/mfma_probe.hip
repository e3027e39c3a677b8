// Measures v_mfma_f32_32x32x16_f16 on gfx950 for the learned-layer
// predictor's product model (tests/golden/make_mfma_fixture.py builds and
// runs it; the outputs are the fixture mfma_f16_32x32x16_gfx950.npz).
// Each wave computes one case: D = A(32x16) * B(16x32) + C(32x32).
// Inputs are generated on the host (deterministic xorshift, six operand
// regimes: plain, wide exponents, cancellation, fp16 subnormals, mixed
// scales, powers of two) and written with the outputs.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <vector>
#include <cmath>

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(64) k_probe(const _Float16* A, const _Float16* B, const float* C,
                                              float* D, int ncase) {
  const int cs = blockIdx.x;
  if (cs >= ncase) return;
  const int l = threadIdx.x;
  const _Float16* a = A + (size_t)cs * 512;   // A[m][k], m<32, k<16
  const _Float16* b = B + (size_t)cs * 512;   // B[k][n], k<16, n<32
  const float* c = C + (size_t)cs * 1024;     // C[m][n]
  float* d = D + (size_t)cs * 1024;
  half8 va, vb;
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * (l / 32) + j;
    va[j] = a[(l % 32) * 16 + k];
    vb[j] = b[k * 32 + (l % 32)];
  }
  float16v acc;
  for (int i = 0; i < 16; ++i) {
    const int m = 8 * (i / 4) + 4 * (l / 32) + (i % 4);
    acc[i] = c[m * 32 + (l % 32)];
  }
  acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(va, vb, acc, 0, 0, 0);
  for (int i = 0; i < 16; ++i) {
    const int m = 8 * (i / 4) + 4 * (l / 32) + (i % 4);
    d[m * 32 + (l % 32)] = acc[i];
  }
}

static uint64_t rs = 88172645463325252ull;
static uint64_t rnd() { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }
static double urand() { return (rnd() >> 11) * (1.0 / 9007199254740992.0); }

int main(int argc, char** argv) {
  const int ncase = 1536;
  const char* out = argc > 1 ? argv[1] : "mfma.bin";
  std::vector<_Float16> A(ncase * 512), B(ncase * 512);
  std::vector<float> C(ncase * 1024), D(ncase * 1024);
  for (int cs = 0; cs < ncase; ++cs) {
    const int regime = cs % 6;
    for (int q = 0; q < 512; ++q) {
      double x, y;
      switch (regime) {
        case 0: x = 2 * urand() - 1; y = 2 * urand() - 1; break;                   // plain
        case 1: x = (2 * urand() - 1) * std::ldexp(1.0, (int)(rnd() % 30) - 15);     // wide exponents
                y = (2 * urand() - 1) * std::ldexp(1.0, (int)(rnd() % 30) - 15); break;
        case 2: x = (rnd() & 1 ? 1 : -1) * (1.0 + urand()); y = (rnd() & 1 ? 1 : -1) * (1.0 + urand()); break;  // cancellation
        case 3: x = (2 * urand() - 1) * 6e-5; y = (2 * urand() - 1) * 1e-3; break;   // subnormal range
        case 4: x = (2 * urand() - 1); y = (2 * urand() - 1) * std::ldexp(1.0, (int)(rnd() % 12)); break;
        default: x = std::ldexp(1.0, (int)(rnd() % 20) - 10) * (rnd() & 1 ? 1 : -1);   // powers of two
                 y = std::ldexp(1.0, (int)(rnd() % 20) - 10) * (1.0 + (rnd() % 4) / 1024.0); break;
      }
      A[cs * 512 + q] = (_Float16)x;
      B[cs * 512 + q] = (_Float16)y;
    }
    for (int q = 0; q < 1024; ++q) {
      double z;
      switch (regime) {
        case 1: z = (2 * urand() - 1) * std::ldexp(1.0, (int)(rnd() % 40) - 20); break;
        case 2: z = (rnd() & 1 ? 1 : -1) * (8.0 + urand()); break;
        case 3: z = (2 * urand() - 1) * 1e-8; break;
        case 5: z = std::ldexp(1.0, (int)(rnd() % 30) - 5) * (rnd() & 1 ? 1 : -1); break;
        default: z = 2 * urand() - 1; break;
      }
      C[cs * 1024 + q] = (float)z;
    }
  }
  _Float16 *dA, *dB; float *dC, *dD;
  if (hipMalloc(&dA, A.size() * 2) || hipMalloc(&dB, B.size() * 2) || hipMalloc(&dC, C.size() * 4) ||
      hipMalloc(&dD, D.size() * 4)) { fprintf(stderr, "alloc failed\n"); return 1; }
  if (hipMemcpy(dA, A.data(), A.size() * 2, hipMemcpyHostToDevice) ||
      hipMemcpy(dB, B.data(), B.size() * 2, hipMemcpyHostToDevice) ||
      hipMemcpy(dC, C.data(), C.size() * 4, hipMemcpyHostToDevice)) { fprintf(stderr, "upload failed\n"); return 1; }
  hipLaunchKernelGGL(k_probe, dim3(ncase), dim3(64), 0, 0, dA, dB, dC, dD, ncase);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) { fprintf(stderr, "kernel: %s\n", hipGetErrorString(e)); return 1; }
  if (hipMemcpy(D.data(), dD, D.size() * 4, hipMemcpyDeviceToHost)) { fprintf(stderr, "download failed\n"); return 1; }
  FILE* f = fopen(out, "wb");
  if (!f) { fprintf(stderr, "cannot write %s\n", out); return 1; }
  fwrite(&ncase, 4, 1, f);
  fwrite(A.data(), 2, A.size(), f);
  fwrite(B.data(), 2, B.size(), f);
  fwrite(C.data(), 4, C.size(), f);
  fwrite(D.data(), 4, D.size(), f);
  fclose(f);
  printf("mfma probe: %d cases written to %s\n", ncase, out);
  return 0;
}

// ceiling_probe.hip -- measures the issue and gather ceilings that the sieve's
// and the SPRP kernel's rooflines are quoted against (bench.py's
// VALU_ISSUE_CEILING, the LDS gather costs of the level-0 layouts).  A
// measurement tool built beside the library; the library does not use it.
//
//   ap26_amd/bin/ceiling_probe [device]
//
// VALU: wave-instructions per CU per nominal 2.4 GHz cycle (wall time) for the
// integer and fp64 instructions the kernels issue, 16 (32-bit) or 8 (64-bit)
// independent chains per lane, at 16 and 32 waves per CU (one or two
// 1024-thread workgroups per CU).
// LDS: CU cycles per 64-lane gather instruction, 16 gathers in flight per
// wave, 16 waves per CU, for the sieve's three access patterns:
//   random16 : 16-byte entries of one 137-entry table at random rows (the
//              level-0 pattern without the rotation)
//   rotated  : 137 rows of 256 bytes; lane j reads the 16-byte column
//              (j + k) mod 16 of a random row at step k (the bank-rotated
//              level 0: the 16 lanes of a pass hit 16 different bank groups)
//   random8  : 8-byte entries of one 211-entry table at random rows (level 1)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define PROBE_CHECK(x)                                                                           \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      fprintf(stderr, "ceiling_probe: %s at line %d: %s\n", #x, __LINE__, hipGetErrorString(e_)); \
      exit(1);                                                                                   \
    }                                                                                            \
  } while (0)

namespace {

constexpr double NOMINAL_HZ = 2.4e9;
constexpr int THREADS = 1024;      /* one workgroup = 16 waves */
constexpr int MAX_BLOCKS_PER_CU = 2;
constexpr int ROWS = 137;          /* rows of the rotated layout = the largest rotated prime */
constexpr int TAB16 = ROWS * 16;   /* uint4 entries: 35,072 bytes */
constexpr int ENTRIES8 = 211;      /* 8-byte entries of the level-1 pattern */

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

#define CHAIN32(ASM)                                                            \
  _Pragma("unroll") for (int j = 0; j < 16; j++) asm volatile(ASM : "+v"(a[j]) : "v"(b), "v"(c));

template <int OP>
__global__ void __launch_bounds__(THREADS) valu32_kernel(uint32_t *out, int iters) {
  uint32_t a[16];
  const uint32_t b = mix(threadIdx.x + 7u), c = mix(blockIdx.x + 11u);
  for (int j = 0; j < 16; j++) a[j] = mix(b + (uint32_t)j);
  for (int it = 0; it < iters; it++) {
    if constexpr (OP == 0) { CHAIN32("v_add_u32 %0, %0, %1") }
    if constexpr (OP == 1) { CHAIN32("v_xor_b32 %0, %0, %1") }
    if constexpr (OP == 2) { CHAIN32("v_add3_u32 %0, %0, %1, %2") }
    if constexpr (OP == 3) { CHAIN32("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x80") }
    if constexpr (OP == 4) { CHAIN32("v_mad_i32_i24 %0, %0, %1, %2") }
    if constexpr (OP == 5) { CHAIN32("v_mul_lo_u32 %0, %0, %1") }
    if constexpr (OP == 6) { CHAIN32("v_mul_hi_u32 %0, %0, %1") }
  }
  uint32_t s = 0;
  for (int j = 0; j < 16; j++) s ^= a[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int OP>
__global__ void __launch_bounds__(THREADS) valu64_kernel(uint32_t *out, int iters) {
  uint64_t a[8];
  double f[8];
  const uint32_t c = mix(blockIdx.x + 3u), d = mix(threadIdx.x + 5u) | 1u;
  const double g = 1.0000001, h = 0.5;
  for (int j = 0; j < 8; j++) {
    a[j] = ((uint64_t)mix(c + (uint32_t)j) << 32) | d;
    f[j] = (double)(j + (int)threadIdx.x);
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if constexpr (OP == 0) asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(a[j]) : "v"(c), "v"(d) : "vcc");
      if constexpr (OP == 1) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(f[j]) : "v"(g), "v"(h));
    }
  }
  uint64_t s = 0;
  for (int j = 0; j < 8; j++) s ^= a[j] ^ (uint64_t)f[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)s ^ (uint32_t)(s >> 32);
}

/* MODE 0: random16, 1: rotated, 2: random8 (see the file comment).  Every
   byte offset is formed on the device below the table's size: row < ROWS
   (or < ENTRIES8), column < 16. */
template <int MODE>
__global__ void __launch_bounds__(THREADS) lds_kernel(uint32_t *out, int iters) {
  __shared__ uint4 t[TAB16];
  for (int i = threadIdx.x; i < TAB16; i += blockDim.x)
    t[i] = make_uint4(mix(4u * i), mix(4u * i + 1u), mix(4u * i + 2u), mix(4u * i + 3u));
  __syncthreads();
  const int lane = threadIdx.x & 63;
  uint32_t x = mix(threadIdx.x + 1024u * blockIdx.x + 12345u);
  uint32_t off[16];
  for (int k = 0; k < 16; k++) {
    x = mix(x + (uint32_t)k);
    if constexpr (MODE == 0) off[k] = (x % ROWS) * 16u;
    if constexpr (MODE == 1) off[k] = (x % ROWS) * 256u + (uint32_t)((lane + k) & 15) * 16u;
    if constexpr (MODE == 2) off[k] = (x % ENTRIES8) * 8u;
  }
  const char *base = (const char *)t;
  uint32_t a0 = 0, a1 = 0;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 16; k++) asm volatile("" : "+v"(off[k])); /* reload every iteration */
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if constexpr (MODE == 2) {
        const uint2 w = *(const uint2 *)(base + off[k]);
        a0 ^= w.x;
        a1 ^= w.y;
      } else {
        const uint4 w = *(const uint4 *)(base + off[k]);
        a0 ^= w.x ^ w.y;
        a1 ^= w.z ^ w.w;
      }
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1;
}

struct Probe {
  uint32_t *d_out = nullptr;
  size_t cap = 0; /* entries of d_out */
  int ncu = 0;
  hipEvent_t e0, e1;

  /* Launch `k` on blocks x THREADS, warm once, time once; returns ms. */
  template <class KernelFn>
  float time(KernelFn k, int blocks, int iters) {
    if ((size_t)blocks * THREADS > cap) {
      fprintf(stderr, "ceiling_probe: grid of %d blocks exceeds the output buffer\n", blocks);
      exit(1);
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(THREADS), 0, 0, d_out, iters / 10 + 1);
    PROBE_CHECK(hipGetLastError());
    PROBE_CHECK(hipDeviceSynchronize());
    PROBE_CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(THREADS), 0, 0, d_out, iters);
    PROBE_CHECK(hipGetLastError());
    PROBE_CHECK(hipEventRecord(e1));
    PROBE_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  }

  template <class KernelFn>
  void valu(const char *name, KernelFn k, int chains) {
    const int iters = 40000;
    for (int per_cu = 1; per_cu <= MAX_BLOCKS_PER_CU; per_cu++) {
      const float ms = time(k, ncu * per_cu, iters);
      const double waves = 16.0 * per_cu;
      const double instr = waves * iters * chains; /* wave-instructions per CU */
      printf("valu %-16s waves/CU %2.0f  %8.3f ms  %.3f wave-instr per CU per nominal cycle  (%.1f T lane-ops/s)\n",
             name, waves, ms, instr / (ms * 1e-3 * NOMINAL_HZ), instr * ncu * 64 / (ms * 1e-3) / 1e12);
      fflush(stdout);
    }
  }

  template <class KernelFn>
  void lds(const char *name, KernelFn k) {
    const int iters = 4000;
    const float ms = time(k, ncu, iters);
    const double gathers = 16.0 * iters * 16; /* wave-instructions per CU */
    printf("lds  %-16s waves/CU 16  %8.3f ms  %.2f CU cycles per gather wave-instruction\n", name, ms,
           ms * 1e-3 * NOMINAL_HZ / gathers);
    fflush(stdout);
  }
};

}  // namespace

int main(int argc, char **argv) {
  const int dev = argc > 1 ? atoi(argv[1]) : 0;
  PROBE_CHECK(hipSetDevice(dev));
  Probe p;
  PROBE_CHECK(hipDeviceGetAttribute(&p.ncu, hipDeviceAttributeMultiprocessorCount, dev));
  p.cap = (size_t)p.ncu * MAX_BLOCKS_PER_CU * THREADS;
  PROBE_CHECK(hipMalloc(&p.d_out, p.cap * sizeof(uint32_t)));
  PROBE_CHECK(hipEventCreate(&p.e0));
  PROBE_CHECK(hipEventCreate(&p.e1));
  printf("device %d: %d CUs; rates at the nominal %.1f GHz\n", dev, p.ncu, NOMINAL_HZ / 1e9);
  p.valu("v_add_u32", valu32_kernel<0>, 16);
  p.valu("v_xor_b32", valu32_kernel<1>, 16);
  p.valu("v_add3_u32", valu32_kernel<2>, 16);
  p.valu("v_bitop3_b32", valu32_kernel<3>, 16);
  p.valu("v_mad_i32_i24", valu32_kernel<4>, 16);
  p.valu("v_mul_lo_u32", valu32_kernel<5>, 16);
  p.valu("v_mul_hi_u32", valu32_kernel<6>, 16);
  p.valu("v_mad_u64_u32", valu64_kernel<0>, 8);
  p.valu("v_fma_f64", valu64_kernel<1>, 8);
  p.lds("random16", lds_kernel<0>);
  p.lds("rotated", lds_kernel<1>);
  p.lds("random8", lds_kernel<2>);
  PROBE_CHECK(hipFree(p.d_out));
  printf("done\n");
  return 0;
}

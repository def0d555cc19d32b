// Test-only probe: the cost of the factor kernel's 16x16 diagonal Cholesky+inverse
// step (diag_chol_inv16) beside different work on the other waves of its workgroup.
// Wave 0 factors one SPD tile from LDS `reps` times; waves 1-7 meanwhile
//   0: idle at the barrier            1: MFMA tile products on LDS operands
//   2: dependent VALU FMAs             3: as 1, but not on wave 0's SIMD (waves 4k idle)
//   4: MFMAs on register operands      5: LDS reads
// Clock stamps per wave; launched with 256 (one workgroup per CU) or 512 workgroups of
// 512 threads with the FP32 kernel's LDS footprint and register cap.
#include "../../blockstructuredsolvers_amd/csrc/chain_factor.hpp"

namespace {
using namespace bss;

template <int MODE>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
    diag_probe_kernel(float* out, unsigned long long* cyc, int reps) {
  __shared__ float Abuf[64 * 68];
  __shared__ float Ubuf[64 * 68];
  __shared__ float R[2 * 64 * 68];
  const int tid = threadIdx.x;
  for (int e = tid; e < 64 * 68; e += 512) {
    const int c = e / 68, r = e % 68;
    Abuf[e] = (r == c) ? 16.f : 0.01f * ((r + c) % 13);
    Ubuf[e] = 0.f;
  }
  for (int e = tid; e < 2 * 64 * 68; e += 512) R[e] = 0.001f * (e % 17);
  __syncthreads();
  const int w = wave_id();
  float acc0 = 0.f;
  const unsigned long long t0 = clock64();
  if (w == 0) {
    for (int rep = 0; rep < reps; ++rep) {
      const bool ok = diag_chol_inv16<float>(Abuf, 68, Ubuf, 68);
      acc0 += ok ? Ubuf[lane_id() & 15] : 1.f;
    }
  } else if (MODE == 1 || (MODE == 3 && (w & 3) != 0)) {
    Mfma<float>::acc_t a = Mfma<float>::zero();
    for (int rep = 0; rep < reps * 4; ++rep)
      tile_mma<float, true, false, false, 64>(a, R + 16 * (w & 3) * 68, 68,
                                              R + (64 + 16 * (rep & 3)) * 68, 68);
    acc0 = a[0] + a[1] + a[2] + a[3];
  } else if (MODE == 2) {
    float x = tid * 1e-3f, y = 1.0001f;
    for (int rep = 0; rep < reps * 600; ++rep) {
      x = fmaf(x, y, 1e-7f);
      y = fmaf(y, x, -1e-7f);
    }
    acc0 = x + y;
  } else if (MODE == 4) {
    Mfma<float>::acc_t a = Mfma<float>::zero();
    const float p = R[tid], q = R[tid + 512];
    for (int rep = 0; rep < reps * 64; ++rep) {
      a = Mfma<float>::mma(p, q, a);
      a = Mfma<float>::mma(q, p, a);
    }
    acc0 = a[0] + a[1] + a[2] + a[3];
  } else if (MODE == 5) {
    float s = 0.f;
    for (int rep = 0; rep < reps * 64; ++rep) {
      const float4 v = *(const float4*)(R + ((rep * 512 + tid * 4) % (2 * 64 * 64)));
      s += v.x + v.y + v.z + v.w;
    }
    acc0 = s;
  }
  const unsigned long long t1 = clock64();
  __syncthreads();
  out[blockIdx.x * 512 + tid] = acc0;
  if ((tid & 63) == 0) {
    unsigned hw;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    cyc[(blockIdx.x * 8 + w) * 2] = t1 - t0;
    cyc[(blockIdx.x * 8 + w) * 2 + 1] = (hw >> 4) & 3;  // SIMD of the wave
  }
}

typedef void (*kfn)(float*, unsigned long long*, int);
kfn pick(int mode) {
  switch (mode) {
    case 0: return diag_probe_kernel<0>;
    case 1: return diag_probe_kernel<1>;
    case 2: return diag_probe_kernel<2>;
    case 3: return diag_probe_kernel<3>;
    case 4: return diag_probe_kernel<4>;
    default: return diag_probe_kernel<5>;
  }
}
}  // namespace

// cyc_host: grid * 8 * 2 values (per wave: cycles of its work, its SIMD)
extern "C" int diag_probe(int mode, int grid, int reps, unsigned long long* cyc_host) {
  if (mode < 0 || mode > 5 || grid < 1 || grid > 4096 || reps < 1 || reps > 1024) return -1;
  float* out = nullptr;
  unsigned long long* cyc = nullptr;
  if (hipMalloc(&out, sizeof(float) * grid * 512) != hipSuccess) return -2;
  if (hipMalloc(&cyc, sizeof(unsigned long long) * grid * 16) != hipSuccess) {
    (void)hipFree(out);
    return -2;
  }
  hipLaunchKernelGGL(pick(mode), dim3(grid), dim3(512), 0, 0, out, cyc, reps);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(cyc_host, cyc, sizeof(unsigned long long) * grid * 16, hipMemcpyDeviceToHost);
  (void)hipFree(out);
  (void)hipFree(cyc);
  return e == hipSuccess ? 0 : -3;
}

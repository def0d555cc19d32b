// tile_probe.hip -- test-only library exercising bss::tile_mma (common.hpp) on one
// 16x16 output tile, so the MFMA operand/result lane layout the solver relies on is
// checked directly against numpy.  Built by __graft_entry__.build(); never linked into
// libbss.
#include <hip/hip_runtime.h>

#include "../../blockstructuredsolvers_amd/csrc/common.hpp"

template <typename T, bool TA, bool TB, bool NEG>
__global__ void __launch_bounds__(64) probe_kernel(const T* X, int ldx, const T* Y, int ldy,
                                                   T* D, int K) {
  typename bss::Mfma<T>::acc_t acc = bss::tile_load<T>(D, 16);
  switch (K) {
    case 16: bss::tile_mma<T, TA, TB, NEG, 16>(acc, X, ldx, Y, ldy); break;
    case 32: bss::tile_mma<T, TA, TB, NEG, 32>(acc, X, ldx, Y, ldy); break;
    case 48: bss::tile_mma<T, TA, TB, NEG, 48>(acc, X, ldx, Y, ldy); break;
    default: bss::tile_mma<T, TA, TB, NEG, 64>(acc, X, ldx, Y, ldy); break;
  }
  bss::tile_store<T>(D, 16, acc);
}

template <typename T>
static int run(int ta, int tb, int neg, const T* X, int ldx, const T* Y, int ldy, T* D, int K) {
#define P(a, b, c)                                                                      \
  if (ta == a && tb == b && neg == c) {                                                 \
    hipLaunchKernelGGL((probe_kernel<T, a, b, c>), dim3(1), dim3(64), 0, 0, X, ldx, Y, ldy, \
                       D, K);                                                           \
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                \
  }
  P(0, 0, 0) P(0, 1, 0) P(1, 0, 0) P(1, 1, 0) P(0, 0, 1) P(0, 1, 1) P(1, 0, 1) P(1, 1, 1)
#undef P
  return -1;
}

// D (16x16, column-major, ld 16) += op(X) op(Y), X/Y device buffers.  Returns 0 on
// success.  dtype: 0 float, 1 double.
extern "C" int tile_probe(int dtype, int ta, int tb, int neg, const void* X, int ldx,
                          const void* Y, int ldy, void* D, int K) {
  if (K <= 0 || K % 16 || K > 64) return -1;
  if (dtype == 1)
    return run<double>(ta, tb, neg, (const double*)X, ldx, (const double*)Y, ldy, (double*)D, K);
  return run<float>(ta, tb, neg, (const float*)X, ldx, (const float*)Y, ldy, (float*)D, K);
}

// Mixed path: FP64 user blocks read and rounded to FP32 on load by the top-level
// factor kernel; reduced diagonals assembled in FP64 from the FP64 separator blocks;
// the FP32 <-> FP64 hand-off kernels used once separator levels run in FP64.
#include "kernels_impl.hpp"
namespace bss {
template hipError_t launch_factor<double, float>(int, int, const FactorLevelArgs&, int, hipStream_t);
template hipError_t launch_assemble_diag<double, float, float>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_assemble_diag<double, float, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_assemble_rhs<float, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_scatter_sep<float, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);

template <typename S, typename D>
hipError_t launch_convert(const ConvRec<S, D>* recs, int nrec, hipStream_t s) {
  if (nrec <= 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<S, D>), dim3(nrec, 64), dim3(256), 0, s, recs);
  return hipGetLastError();
}
template hipError_t launch_convert<float, double>(const ConvRec<float, double>*, int, hipStream_t);
template hipError_t launch_convert<double, float>(const ConvRec<double, float>*, int, hipStream_t);
}  // namespace bss

// FP64 plain path: level kernels in double.
#include "kernels_impl.hpp"
namespace bss {
template hipError_t launch_factor<double, double>(int, int, const FactorLevelArgs&, int, hipStream_t);
template hipError_t launch_forward<double>(int, const SolveLevelArgs&, int, hipStream_t);
template hipError_t launch_backward<double>(int, const SolveLevelArgs&, int, hipStream_t);
template hipError_t launch_assemble_diag<double, double, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_assemble_rhs<double, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_scatter_sep<double, double>(const SepRec*, const LevelSys*, int, int, hipStream_t);
}  // namespace bss

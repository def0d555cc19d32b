// FP32 path: level kernels in float (plain FP32 solves and the factor/solve of the
// mixed-precision path below its FP64 top-level input).
#include "kernels_impl.hpp"
namespace bss {
template hipError_t launch_factor<float, float>(int, int, const FactorLevelArgs&, int, hipStream_t);
template hipError_t launch_forward<float>(int, const SolveLevelArgs&, int, hipStream_t);
template hipError_t launch_backward<float>(int, const SolveLevelArgs&, int, hipStream_t);
template hipError_t launch_assemble_diag<float, float, float>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_assemble_rhs<float, float>(const SepRec*, const LevelSys*, int, int, hipStream_t);
template hipError_t launch_scatter_sep<float, float>(const SepRec*, const LevelSys*, int, int, hipStream_t);
}  // namespace bss

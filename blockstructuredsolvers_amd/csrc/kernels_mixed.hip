// Mixed path: FP64 user blocks read and rounded to FP32 on load by the top-level
// factor kernel; reduced diagonals assembled in FP64 from the FP64 separator blocks.
#include "kernels_impl.hpp"
namespace bss {
template hipError_t launch_factor<double, float>(int, int, const FactorLevelArgs&, int, hipStream_t);
template hipError_t launch_assemble_diag<double, float>(const SepRec*, const LevelSys*, int, int, hipStream_t);
}  // namespace bss

// FP64 plain path: level kernels in double.
#include "kernels_impl.hpp"
namespace bss {
template hipError_t launch_factor<double, double>(int, int, const FactorLevelArgs<double>&, int, int, hipStream_t);
template hipError_t launch_forward<double>(int, const SolveLevelArgs<double>&, int, int, hipStream_t);
template hipError_t launch_backward<double>(int, const SolveLevelArgs<double>&, int, int, hipStream_t);
template hipError_t launch_assemble_diag<double, double>(const double*, const double*, const double*, const int*, double*, int, int, int, long long, long long, long long, hipStream_t);
template hipError_t launch_assemble_rhs<double>(const double*, const double*, const double*, const int*, double*, int, int, int, long long, long long, long long, hipStream_t);
template hipError_t launch_scatter_sep<double>(const double*, const int*, double*, int, int, int, long long, long long, hipStream_t);
}  // namespace bss

// Instantiation of the factored-path kernels for __bf16 operands split in 3
// part(s) (see gsc_gemm.h); one file per operand format so the builds run in
// parallel.
#include "gsc_gemm.h"

namespace gsc {
template hipError_t tables_launch<__bf16, 3>(const TableArgs&, dim3, hipStream_t);
template hipError_t factored_gemm_launch<__bf16, 3>(const GemmArgs&, dim3, hipStream_t);
template hipError_t ens_launch<__bf16, 3>(const EnsLaunch&, int, int, bool, hipStream_t);
}  // namespace gsc

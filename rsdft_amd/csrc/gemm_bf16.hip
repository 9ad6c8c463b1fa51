// BF16 hi+lo split GEMMs on MFMA (placeholder until the kernels land).
#include "common.h"

namespace rsdft {

template <class T>
int split_gemm_tn(const float *, const float *, int, long, double, double *)
{
    return set_error(kErrState, "BF16 split GEMM not built yet");
}

template <class T>
int split_gemm_nn(const float *, const double *, int, long, float *)
{
    return set_error(kErrState, "BF16 split GEMM not built yet");
}

template int split_gemm_tn<float>(const float *, const float *, int, long, double, double *);
template int split_gemm_tn<double>(const float *, const float *, int, long, double, double *);
template int split_gemm_nn<float>(const float *, const double *, int, long, float *);
template int split_gemm_nn<double>(const float *, const double *, int, long, float *);

}  // namespace rsdft

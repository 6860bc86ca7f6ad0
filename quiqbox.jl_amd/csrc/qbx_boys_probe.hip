// qbx_boys_probe.hip -- the ERI kernels' Boys function evaluated on its own
// (qbx_boys_device), for tests: the same LDS staging of the order-L rows and
// the same boys_lds<L> routine the class kernels inline.
#include "qbx_eri.h"

namespace qbx {
namespace {

template <int L>
__global__ __launch_bounds__(256) void boys_probe_kernel(const double* __restrict__ rows, const double* __restrict__ T,
                                                         int n, double* __restrict__ F)
{
    __shared__ double2 sboys[BOYS_HEAD2 + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE / 2];
    stage_boys<256>(sboys, rows);
    __syncthreads();
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= n) return;
    double f[L + 1];
    boys_lds<L>(T[i], f, (const double*)sboys + BOYS_HEAD);
#pragma unroll
    for (int m = 0; m <= L; ++m) F[(size_t)i * (L + 1) + m] = f[m];
}

template <int L> void launch(const double* rows, const double* T, int n, double* F, hipStream_t s)
{
    hipLaunchKernelGGL(boys_probe_kernel<L>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, T, n, F);
}

}  // namespace

void boys_probe(const double* rows, int L, const double* T, int n, double* F, hipStream_t s)
{
    switch (L) {
    case 0: launch<0>(rows, T, n, F, s); break;
    case 1: launch<1>(rows, T, n, F, s); break;
    case 2: launch<2>(rows, T, n, F, s); break;
    case 3: launch<3>(rows, T, n, F, s); break;
    case 4: launch<4>(rows, T, n, F, s); break;
    case 5: launch<5>(rows, T, n, F, s); break;
    case 6: launch<6>(rows, T, n, F, s); break;
    case 7: launch<7>(rows, T, n, F, s); break;
    case 8: launch<8>(rows, T, n, F, s); break;
    default: throw QbxError(QBX_EINVAL, "Boys order must be in [0, 8]");
    }
    QBX_HIP(hipGetLastError());
}

}  // namespace qbx

// CSR operands of the request server, expanded to dense row-major FP32 in
// HBM before the dense compress path.  Sketch cost is dominated by the
// per-k transforms (2 d n b log b), not by reading the operand, so a 1%
// dense CSR row costs the same transform work either way; densifying costs
// one HBM write of the operand.
#include "kernels.hpp"

namespace cm {

namespace {

// One workgroup per row: zero the row, then add its nonzeros (a repeated
// column is summed, as a CSR matrix means).  Out-of-range columns are
// skipped here; the host rejects them before launch.
__global__ void __launch_bounds__(256) k_csr_dense(const int32_t* __restrict__ indptr,
                                                   const int32_t* __restrict__ indices,
                                                   const float* __restrict__ values, int cols,
                                                   int64_t ld, float* __restrict__ out) {
  const int r = blockIdx.x;
  float* row = out + int64_t(r) * ld;
  for (int c = threadIdx.x; c < cols; c += 256) row[c] = 0.f;
  __syncthreads();
  const int a = indptr[r], e = indptr[r + 1];
  for (int k = a + threadIdx.x; k < e; k += 256) {
    const int c = indices[k];
    if (c >= 0 && c < cols) atomicAdd(&row[c], values[k]);
  }
}

}  // namespace

hipError_t launch_csr_dense(const int32_t* indptr, const int32_t* indices, const float* values,
                            int rows, int cols, int64_t ld, float* out, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  k_csr_dense<<<dim3(unsigned(rows)), dim3(256), 0, s>>>(indptr, indices, values, cols, ld, out);
  return hipGetLastError();
}

}  // namespace cm

// Dense matrix copy between device buffers, optionally transposed:
// B (m x n, ld ldb) = A or A^T.  This is the device side of kernels::assign
// between dense hip_tag matrices (include/remora/kernels/hip/matrix_assign.hpp),
// which a matrix assignment between row- and column-major storage needs.
//
// The plain copy is a strided 2-D DMA copy.  The transposed copy stages a
// 64 x 64 tile through LDS (rows padded to 65 elements, so the column reads
// of the write phase do not conflict): both the reads of A and the writes of
// B run along 64 consecutive elements of a column.  Only the m x n elements
// of each matrix are touched.
#include "rmh_context.hpp"

namespace rmh {
namespace {

constexpr int kTile = 64;

template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(int64_t m, int64_t n, const T* A, int64_t lda, T* B,
                                                        int64_t ldb) {
    __shared__ T tile[kTile][kTile + 1];
    // B tile origin: rows [r0, r0+64), columns [c0, c0+64); B(r, c) = A(c, r)
    const int64_t r0 = (int64_t)blockIdx.x * kTile, c0 = (int64_t)blockIdx.y * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;  // 64 x 4
    // read A rows [c0, c0+64) (lane tx runs down a column of A), columns [r0, r0+64)
    for (int k = ty; k < kTile; k += 4) {
        const int64_t ai = c0 + tx, aj = r0 + k;
        if (ai < n && aj < m) tile[k][tx] = A[ai + aj * lda];
    }
    __syncthreads();
    for (int k = ty; k < kTile; k += 4) {
        const int64_t bi = r0 + tx, bj = c0 + k;
        if (bi < m && bj < n) B[bi + bj * ldb] = tile[tx][k];
    }
}

}  // namespace

template <typename T>
rmh_status copy_matrix_impl(bool trans, int64_t m, int64_t n, const T* A, int64_t lda, T* B, int64_t ldb,
                            hipStream_t s) {
    if (m == 0 || n == 0) return RMH_STATUS_SUCCESS;
    if (!trans) {
        RMH_TRY_HIP(hipMemcpy2DAsync(B, (size_t)ldb * sizeof(T), A, (size_t)lda * sizeof(T), (size_t)m * sizeof(T),
                                     (size_t)n, hipMemcpyDeviceToDevice, s));
        return RMH_STATUS_SUCCESS;
    }
    const int64_t gx = (m + kTile - 1) / kTile, gy = (n + kTile - 1) / kTile;
    if (gx > 0x7fffffff || gy > 65535) return fail(RMH_STATUS_INVALID_VALUE, "copy_matrix: matrix too large");
    hipLaunchKernelGGL(transpose_kernel<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, m, n, A, lda, B, ldb);
    return check_launch("transpose_kernel");
}

template rmh_status copy_matrix_impl<double>(bool, int64_t, int64_t, const double*, int64_t, double*, int64_t,
                                             hipStream_t);
template rmh_status copy_matrix_impl<float>(bool, int64_t, int64_t, const float*, int64_t, float*, int64_t,
                                            hipStream_t);

}  // namespace rmh

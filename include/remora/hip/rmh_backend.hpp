// remora/hip/rmh_backend.hpp -- hip::blas_device and get_blas on the rmh_*
// C-ABI: the class the reference's hip_tag kernel glue calls
// (include/remora/kernels/hip/{gemm,gemv,syrk,trmm,trmv,trsm,trsv}.hpp), with
// the method set and argument order of the cuBLAS-backed original
// (include/remora/hip/cublas_backend.hpp:86-450): bool operation flags,
// std::size_t sizes and leading dimensions, scalars by value, column-major
// device pointers, and the stream to queue on.  Each method forwards to the
// matching rmh_s* / rmh_d* entry point and throws hip::rmh_exception on a
// non-success status (the original threw cublas_exception).
//
// Differences from the original, by design:
//   - sizes stay 64-bit (the original narrowed them to int for cuBLAS,
//     cublas_backend.hpp:136-141);
//   - the library handle is the thread's per-device rmh_handle of
//     get_handle(), so a blas_device owns nothing and moving it is trivial.
#ifndef REMORA_HIP_RMH_BACKEND_HPP
#define REMORA_HIP_RMH_BACKEND_HPP

#include <cstddef>
#include <utility>
#include <vector>

#include "abi.hpp"
#include "core.hpp"

namespace remora {
namespace hip {

class blas_device {
public:
    explicit blas_device(device& d) : m_device(&d) {}
    blas_device(blas_device const&) = delete;
    blas_device& operator=(blas_device const&) = delete;
    blas_device(blas_device&&) noexcept = default;
    blas_device& operator=(blas_device&&) noexcept = default;

    // C = alpha op(A) op(B) + beta C (cublas_backend.hpp:121-166)
    void gemm(bool transA, bool transB, std::size_t m, std::size_t n, std::size_t k, float alpha, float const* A,
              std::size_t ldA, float const* B, std::size_t ldB, float beta, float* C, std::size_t ldC,
              stream const& s) {
        check_rmh(abi::gemm(prepare(), transA, transB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, s.abi()));
    }
    void gemm(bool transA, bool transB, std::size_t m, std::size_t n, std::size_t k, double alpha, double const* A,
              std::size_t ldA, double const* B, std::size_t ldB, double beta, double* C, std::size_t ldC,
              stream const& s) {
        check_rmh(abi::gemm(prepare(), transA, transB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, s.abi()));
    }

    // v = alpha op(A) x + beta v (cublas_backend.hpp:170-213)
    void gemv(bool transA, std::size_t m, std::size_t n, float alpha, float const* A, std::size_t ldA,
              float const* x, std::size_t stridex, float beta, float* v, std::size_t stridev, stream const& s) {
        check_rmh(abi::gemv(prepare(), transA, m, n, alpha, A, ldA, x, stridex, beta, v, stridev, s.abi()));
    }
    void gemv(bool transA, std::size_t m, std::size_t n, double alpha, double const* A, std::size_t ldA,
              double const* x, std::size_t stridex, double beta, double* v, std::size_t stridev, stream const& s) {
        check_rmh(abi::gemv(prepare(), transA, m, n, alpha, A, ldA, x, stridex, beta, v, stridev, s.abi()));
    }

    // B = alpha op(A) B (left) or alpha B op(A), A triangular (cublas_backend.hpp:217-263)
    void trmm(bool leftA, bool upperA, bool transA, bool unitA, std::size_t m, std::size_t n, float alpha,
              float const* A, std::size_t ldA, float* B, std::size_t ldB, stream const& s) {
        check_rmh(abi::trmm(prepare(), leftA, upperA, transA, unitA, m, n, alpha, A, ldA, B, ldB, s.abi()));
    }
    void trmm(bool leftA, bool upperA, bool transA, bool unitA, std::size_t m, std::size_t n, double alpha,
              double const* A, std::size_t ldA, double* B, std::size_t ldB, stream const& s) {
        check_rmh(abi::trmm(prepare(), leftA, upperA, transA, unitA, m, n, alpha, A, ldA, B, ldB, s.abi()));
    }

    // v = op(A) v, A triangular (cublas_backend.hpp:265-299)
    void trmv(bool upperA, bool transA, bool unitA, std::size_t n, float const* A, std::size_t ldA, float* v,
              std::size_t stridev, stream const& s) {
        check_rmh(abi::trmv(prepare(), upperA, transA, unitA, n, A, ldA, v, stridev, s.abi()));
    }
    void trmv(bool upperA, bool transA, bool unitA, std::size_t n, double const* A, std::size_t ldA, double* v,
              std::size_t stridev, stream const& s) {
        check_rmh(abi::trmv(prepare(), upperA, transA, unitA, n, A, ldA, v, stridev, s.abi()));
    }

    // op(A) X = alpha B (left) or X op(A) = alpha B, in place in B (cublas_backend.hpp:303-344)
    void trsm(bool leftA, bool upperA, bool transA, bool unitA, std::size_t m, std::size_t n, float alpha,
              float const* A, std::size_t ldA, float* B, std::size_t ldB, stream const& s) {
        check_rmh(abi::trsm(prepare(), leftA, upperA, transA, unitA, m, n, alpha, A, ldA, B, ldB, s.abi()));
    }
    void trsm(bool leftA, bool upperA, bool transA, bool unitA, std::size_t m, std::size_t n, double alpha,
              double const* A, std::size_t ldA, double* B, std::size_t ldB, stream const& s) {
        check_rmh(abi::trsm(prepare(), leftA, upperA, transA, unitA, m, n, alpha, A, ldA, B, ldB, s.abi()));
    }

    // op(A) x = v in place in v (cublas_backend.hpp:348-383)
    void trsv(bool upperA, bool transA, bool unitA, std::size_t n, float const* A, std::size_t ldA, float* v,
              std::size_t stridev, stream const& s) {
        check_rmh(abi::trsv(prepare(), upperA, transA, unitA, n, A, ldA, v, stridev, s.abi()));
    }
    void trsv(bool upperA, bool transA, bool unitA, std::size_t n, double const* A, std::size_t ldA, double* v,
              std::size_t stridev, stream const& s) {
        check_rmh(abi::trsv(prepare(), upperA, transA, unitA, n, A, ldA, v, stridev, s.abi()));
    }

    // C = alpha op(A) op(A)^T + beta C on one triangle (cublas_backend.hpp:389-431)
    void syrk(bool upperA, bool transA, std::size_t n, std::size_t k, float alpha, float const* A, std::size_t ldA,
              float beta, float* C, std::size_t ldC, stream const& s) {
        check_rmh(abi::syrk(prepare(), upperA, transA, n, k, alpha, A, ldA, beta, C, ldC, s.abi()));
    }
    void syrk(bool upperA, bool transA, std::size_t n, std::size_t k, double alpha, double const* A,
              std::size_t ldA, double beta, double* C, std::size_t ldC, stream const& s) {
        check_rmh(abi::syrk(prepare(), upperA, transA, n, k, alpha, A, ldA, beta, C, ldC, s.abi()));
    }

    device& queue() const { return *m_device; }

private:
    rmh_handle prepare() const {
        m_device->set_device();
        return get_handle(*m_device);
    }
    device* m_device;
};

// One blas_device per thread per device (cublas_backend.hpp:442-450).  The
// cache lives until the process ends, like get_stream's.
inline blas_device& get_blas(device& d) {
    thread_local std::vector<std::pair<int, blas_device*>>* cache = new std::vector<std::pair<int, blas_device*>>();
    for (auto& e : *cache)
        if (e.first == d.device_id()) return *e.second;
    cache->emplace_back(d.device_id(), new blas_device(d));
    return *cache->back().second;
}

}  // namespace hip
}  // namespace remora

#endif

// remora/hip/dense.hpp -- dense matrices and vectors of the hip_tag
// front-end (and the minimal cpu_tag matrix the transport needs).
//
// Mirrors (same names and meaning, own code):
//   matrix_expression / matrix_container / vector_expression CRTP bases ... include/remora/expression_types.hpp:45-115
//   dense_matrix_storage{values, leading_dimension} ....................... include/remora/detail/storage.hpp:71-91
//   matrix<T, Orientation, hip_tag> ........................................ include/remora/hip/dense.hpp:451-656
//   dense_matrix_adaptor (views), trans(), subrange() ...................... include/remora/proxy_expressions.hpp:88-99, 210-240
//   copy_to_device / copy_to_cpu ........................................... include/remora/hip/copy.hpp:468-502
// Storage: row_major keeps element (i,j) at values[i*ld + j], column_major
// at values[i + j*ld] (detail/storage.hpp, detail/structure.hpp:96,141).
#ifndef REMORA_HIP_DENSE_HPP
#define REMORA_HIP_DENSE_HPP

#include <algorithm>
#include <cassert>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "abi.hpp"
#include "core.hpp"

namespace remora {

template <class E, class Device>
struct matrix_expression {
    typedef Device device_type;
    E& operator()() { return *static_cast<E*>(this); }
    E const& operator()() const { return *static_cast<E const*>(this); }
};
template <class C, class Device>
struct matrix_container : public matrix_expression<C, Device> {
    C& operator()() { return *static_cast<C*>(this); }
    C const& operator()() const { return *static_cast<C const*>(this); }
};
template <class E, class Device>
struct vector_expression {
    typedef Device device_type;
    E& operator()() { return *static_cast<E*>(this); }
    E const& operator()() const { return *static_cast<E const*>(this); }
};

// Expressions evaluated into a target by their own kernels (products,
// solves): they give size, queue and assign_to / plus_assign_to, the way the
// reference's blockwise expressions do (detail/matrix_expression_classes.hpp,
// solve.hpp).
template <class E>
struct matrix_lazy_expression {
    E const& operator()() const { return *static_cast<E const*>(this); }
};
template <class E>
struct vector_lazy_expression {
    E const& operator()() const { return *static_cast<E const*>(this); }
};

template <class T>
struct dense_matrix_storage {
    T* values;
    std::size_t leading_dimension;
};
template <class T>
struct dense_vector_storage {
    T* values;
    std::size_t stride;
};

namespace detail {
template <class O>
inline std::size_t element_index(std::size_t i, std::size_t j, std::size_t ld) {
    return O::is_row_major ? i * ld + j : i + j * ld;
}
}  // namespace detail

// ---------------------------------------------------------------- views --
// Non-owning view of a dense device matrix (the role of
// dense_matrix_adaptor<T, Orientation, hip_tag>).
template <class T, class Orientation, class Device>
class matrix_view;

template <class T, class Orientation>
class matrix_view<T, Orientation, hip_tag> : public matrix_expression<matrix_view<T, Orientation, hip_tag>, hip_tag> {
public:
    typedef T value_type;
    typedef Orientation orientation;
    typedef dense_matrix_storage<T> storage_type;
    matrix_view(T* values, std::size_t size1, std::size_t size2, std::size_t ld, hip::device& q)
        : m_values(values), m_size1(size1), m_size2(size2), m_ld(ld), m_queue(&q) {}
    std::size_t size1() const { return m_size1; }
    std::size_t size2() const { return m_size2; }
    storage_type raw_storage() const { return {m_values, m_ld}; }
    hip::device& queue() const { return *m_queue; }

private:
    T* m_values;
    std::size_t m_size1, m_size2, m_ld;
    hip::device* m_queue;
};

namespace detail {
// dst = src for dense device matrices of the same size, any orientations, on
// the device: the same layout is a 2-D copy of the lines (rows for
// row_major, columns otherwise), a different layout a transposed copy
// (rmh_?copy_matrix); only the size1 x size2 elements of src are read and of
// dst written.
template <class Dst, class Src>
void copy_dense(Dst const& dst_m, Src const& src_m) {
    typedef typename Dst::orientation O;
    if (dst_m.size1() != src_m.size1() || dst_m.size2() != src_m.size2())
        throw std::invalid_argument("[assign] sizes do not match");
    if (dst_m.size1() == 0 || dst_m.size2() == 0) return;
    auto src = src_m.raw_storage();
    auto dst = dst_m.raw_storage();
    hip::device& q = dst_m.queue();
    q.set_device();
    // the destination's memory as a column-major matrix
    const std::size_t rows = O::is_row_major ? dst_m.size2() : dst_m.size1();
    const std::size_t cols = O::is_row_major ? dst_m.size1() : dst_m.size2();
    const bool transposed = !std::is_same<typename Src::orientation, O>::value;
    hip::check_rmh(hip::abi::copy_matrix(hip::get_handle(q), transposed, rows, cols, src.values,
                                         src.leading_dimension, dst.values, dst.leading_dimension,
                                         hip::get_stream(q).abi()));
}
}  // namespace detail

// ---------------------------------------------------- device container --
template <class T, class Orientation = row_major, class Device = cpu_tag>
class matrix;

template <class T, class Orientation>
class matrix<T, Orientation, hip_tag> : public matrix_container<matrix<T, Orientation, hip_tag>, hip_tag> {
public:
    typedef T value_type;
    typedef Orientation orientation;
    typedef dense_matrix_storage<T> storage_type;
    typedef hip_tag device_type;

    explicit matrix(hip::device& q = hip::device_manager::instance().default_device()) : m_queue(&q) {}
    matrix(std::size_t size1, std::size_t size2,
           hip::device& q = hip::device_manager::instance().default_device())
        : m_queue(&q), m_size1(size1), m_size2(size2), m_data(size1 * size2, q) {}
    matrix(matrix const& o) : matrix(o.size1(), o.size2(), o.queue()) { assign_from(o); }
    matrix(matrix&&) noexcept = default;
    template <class E>
    matrix(matrix_expression<E, hip_tag> const& e) : matrix(e().size1(), e().size2(), e().queue()) {
        assign_from(e());
    }
    matrix& operator=(matrix const& o) {
        resize(o.size1(), o.size2());
        assign_from(o);
        return *this;
    }
    matrix& operator=(matrix&&) noexcept = default;
    template <class E>
    matrix& operator=(matrix_expression<E, hip_tag> const& e) {
        resize(e().size1(), e().size2());
        assign_from(e());
        return *this;
    }
    template <class E>
    matrix(matrix_lazy_expression<E> const& e) : matrix(e().size1(), e().size2(), e().queue()) {
        e().assign_to(*this);
    }
    template <class E>
    matrix& operator=(matrix_lazy_expression<E> const& e) {
        // a product or solve may read this matrix: evaluate into a temporary
        matrix tmp(e);
        *this = std::move(tmp);
        return *this;
    }

    std::size_t size1() const { return m_size1; }
    std::size_t size2() const { return m_size2; }
    std::size_t leading_dimension() const { return std::max<std::size_t>(1, Orientation::is_row_major ? m_size2 : m_size1); }
    storage_type raw_storage() const { return {m_data.get(), leading_dimension()}; }
    hip::device& queue() const { return *m_queue; }
    void resize(std::size_t s1, std::size_t s2) {
        if (s1 * s2 != m_data.size()) m_data = hip::buffer<T>(s1 * s2, *m_queue);
        m_size1 = s1;
        m_size2 = s2;
    }

private:
    template <class E>
    void assign_from(E const& e) {
        detail::copy_dense(*this, e);
    }

    hip::device* m_queue;
    std::size_t m_size1 = 0, m_size2 = 0;
    hip::buffer<T> m_data;
};

// Non-owning view of a dense device vector with a stride (the role of
// dense_vector_adaptor<T, hip_tag>).
template <class T, class Device>
class vector_view;

template <class T>
class vector_view<T, hip_tag> : public vector_expression<vector_view<T, hip_tag>, hip_tag> {
public:
    typedef T value_type;
    typedef dense_vector_storage<T> storage_type;
    vector_view(T* values, std::size_t size, std::size_t stride, hip::device& q)
        : m_values(values), m_size(size), m_stride(stride), m_queue(&q) {}
    std::size_t size() const { return m_size; }
    storage_type raw_storage() const { return {m_values, m_stride}; }
    hip::device& queue() const { return *m_queue; }

private:
    T* m_values;
    std::size_t m_size, m_stride;
    hip::device* m_queue;
};

namespace detail {
// dst[i * sd] = src[i * ss] for i < n on the thread's stream (a 2-D copy of
// n one-element rows, so any strides work)
template <class T>
void copy_vector(T* dst, std::size_t sd, const T* src, std::size_t ss, std::size_t n, hip::device& q) {
    if (n == 0) return;
    q.set_device();
    hip::check_hip(hipMemcpy2DAsync(dst, sd * sizeof(T), src, ss * sizeof(T), sizeof(T), n, hipMemcpyDeviceToDevice,
                                    hip::get_stream(q).handle()),
                   "hipMemcpy2DAsync");
}
}  // namespace detail

// device vector (contiguous)
template <class T, class Device = cpu_tag>
class vector;

template <class T>
class vector<T, hip_tag> : public vector_expression<vector<T, hip_tag>, hip_tag> {
public:
    typedef T value_type;
    typedef dense_vector_storage<T> storage_type;
    typedef hip_tag device_type;
    explicit vector(std::size_t size = 0, hip::device& q = hip::device_manager::instance().default_device())
        : m_queue(&q), m_data(size, q) {}
    // every element set to init (the reference's vector(size, value))
    vector(std::size_t size, T init, hip::device& q = hip::device_manager::instance().default_device())
        : m_queue(&q), m_data(size, q) {
        if (size == 0) return;
        std::vector<T> h(size, init);
        q.set_device();
        hip::stream& s = hip::get_stream(q);
        hip::check_hip(hipMemcpyAsync(m_data.get(), h.data(), size * sizeof(T), hipMemcpyHostToDevice, s.handle()),
                       "hipMemcpyAsync");
        s.synchronize();
    }
    vector(vector const& o) : vector(o.size(), o.queue()) { detail::copy_vector(m_data.get(), 1, o.m_data.get(), 1, size(), *m_queue); }
    vector(vector&&) noexcept = default;
    template <class E>
    vector(vector_expression<E, hip_tag> const& e) : vector(e().size(), e().queue()) {
        auto src = e().raw_storage();
        detail::copy_vector(m_data.get(), 1, src.values, src.stride, size(), *m_queue);
    }
    template <class E>
    vector(vector_lazy_expression<E> const& e) : vector(e().size(), e().queue()) {
        e().assign_to(*this);
    }
    vector& operator=(vector const& o) {
        resize(o.size());
        detail::copy_vector(m_data.get(), 1, o.m_data.get(), 1, size(), *m_queue);
        return *this;
    }
    vector& operator=(vector&&) noexcept = default;
    template <class E>
    vector& operator=(vector_expression<E, hip_tag> const& e) {
        auto src = e().raw_storage();
        resize(e().size());
        detail::copy_vector(m_data.get(), 1, src.values, src.stride, size(), *m_queue);
        return *this;
    }
    template <class E>
    vector& operator=(vector_lazy_expression<E> const& e) {
        vector tmp(e);  // the expression may read this vector
        *this = std::move(tmp);
        return *this;
    }
    std::size_t size() const { return m_data.size(); }
    storage_type raw_storage() const { return {m_data.get(), 1}; }
    hip::device& queue() const { return *m_queue; }
    void resize(std::size_t n) {
        if (n != m_data.size()) m_data = hip::buffer<T>(n, *m_queue);
    }

private:
    hip::device* m_queue;
    hip::buffer<T> m_data;
};

// ------------------------------------------------------------ proxies --
template <class M>
matrix_view<typename M::value_type, typename M::orientation::transposed_orientation, hip_tag> trans(
    matrix_expression<M, hip_tag> const& m) {
    auto s = m().raw_storage();
    return {s.values, m().size2(), m().size1(), s.leading_dimension, m().queue()};
}

template <class M>
matrix_view<typename M::value_type, typename M::orientation, hip_tag> subrange(matrix_expression<M, hip_tag> const& m,
                                                                             std::size_t r0, std::size_t r1,
                                                                             std::size_t c0, std::size_t c1) {
    assert(r0 <= r1 && r1 <= m().size1() && c0 <= c1 && c1 <= m().size2());
    auto s = m().raw_storage();
    return {s.values + detail::element_index<typename M::orientation>(r0, c0, s.leading_dimension), r1 - r0, c1 - c0,
            s.leading_dimension, m().queue()};
}

// ------------------------------------------------------- cpu container --
// Minimal host matrix used by the transport and the tests.
template <class T, class Orientation>
class matrix<T, Orientation, cpu_tag> {
public:
    typedef T value_type;
    typedef Orientation orientation;
    matrix() = default;
    matrix(std::size_t s1, std::size_t s2, T init = T()) : m_size1(s1), m_size2(s2), m_data(s1 * s2, init) {}
    std::size_t size1() const { return m_size1; }
    std::size_t size2() const { return m_size2; }
    std::size_t leading_dimension() const { return std::max<std::size_t>(1, Orientation::is_row_major ? m_size2 : m_size1); }
    T& operator()(std::size_t i, std::size_t j) {
        return m_data[detail::element_index<Orientation>(i, j, leading_dimension())];
    }
    T const& operator()(std::size_t i, std::size_t j) const {
        return m_data[detail::element_index<Orientation>(i, j, leading_dimension())];
    }
    T* data() { return m_data.data(); }
    T const* data() const { return m_data.data(); }

private:
    std::size_t m_size1 = 0, m_size2 = 0;
    std::vector<T> m_data;
};

// Minimal host vector used by the transport and the tests.
template <class T>
class vector<T, cpu_tag> {
public:
    typedef T value_type;
    explicit vector(std::size_t n = 0, T init = T()) : m_data(n, init) {}
    std::size_t size() const { return m_data.size(); }
    T& operator()(std::size_t i) { return m_data[i]; }
    T const& operator()(std::size_t i) const { return m_data[i]; }
    T* data() { return m_data.data(); }
    T const* data() const { return m_data.data(); }

private:
    std::vector<T> m_data;
};

// ----------------------------------------------------------- transport --
// Host <-> device copies with the layout preserved (one contiguous copy on
// the thread's stream, then a synchronisation; the reference used a
// zero-copy kernel over registered host memory, hip/copy.hpp:362-386).
//
// Downloads of 16 MB and more go through two pinned 64 MB staging buffers:
// the DMA of chunk k+1 runs while the CPU copies chunk k out, which measured
// 27 GB/s against 17 GB/s for a copy into pageable memory on MI355X (uploads
// measured the same either way, 27-28 GB/s, and stay direct).  The buffers
// are allocated on first use and kept for the life of the process; a mutex
// serialises their use across threads.
namespace hip {
namespace detail {

struct download_staging {
    static constexpr std::size_t chunk = std::size_t(64) << 20;
    std::mutex lock;
    void* pinned[2] = {nullptr, nullptr};
    static download_staging& instance() {
        static download_staging* st = new download_staging();  // kept until the process ends
        return *st;
    }
};

inline void download(void* dst, const void* src, std::size_t bytes, stream& s) {
    if (bytes < (std::size_t(16) << 20)) {
        check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s.handle()), "hipMemcpyAsync");
        s.synchronize();
        return;
    }
    download_staging& st = download_staging::instance();
    std::lock_guard<std::mutex> guard(st.lock);
    for (int b = 0; b < 2; ++b)
        if (st.pinned[b] == nullptr)
            check_hip(hipHostMalloc(&st.pinned[b], download_staging::chunk, hipHostMallocDefault), "hipHostMalloc");
    hipEvent_t ev[2] = {nullptr, nullptr};
    check_hip(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), "hipEventCreateWithFlags");
    check_hip(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), "hipEventCreateWithFlags");
    const std::size_t nch = (bytes + download_staging::chunk - 1) / download_staging::chunk;
    auto len = [&](std::size_t k) { return std::min(download_staging::chunk, bytes - k * download_staging::chunk); };
    auto issue = [&](std::size_t k) {
        check_hip(hipMemcpyAsync(st.pinned[k & 1], static_cast<const char*>(src) + k * download_staging::chunk,
                                 len(k), hipMemcpyDeviceToHost, s.handle()),
                  "hipMemcpyAsync");
        check_hip(hipEventRecord(ev[k & 1], s.handle()), "hipEventRecord");
    };
    try {
        issue(0);
        for (std::size_t k = 0; k < nch; ++k) {
            if (k + 1 < nch) issue(k + 1);
            check_hip(hipEventSynchronize(ev[k & 1]), "hipEventSynchronize");
            std::memcpy(static_cast<char*>(dst) + k * download_staging::chunk, st.pinned[k & 1], len(k));
        }
    } catch (...) {
        (void)hipStreamSynchronize(s.handle());
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    check_hip(hipEventDestroy(ev[0]), "hipEventDestroy");
    check_hip(hipEventDestroy(ev[1]), "hipEventDestroy");
}

}  // namespace detail
}  // namespace hip

template <class T, class O>
matrix<T, O, hip_tag> copy_to_device(matrix<T, O, cpu_tag> const& m, hip_tag,
                                     hip::device& q = hip::device_manager::instance().default_device()) {
    matrix<T, O, hip_tag> d(m.size1(), m.size2(), q);
    if (m.size1() * m.size2() == 0) return d;
    hip::stream& s = hip::get_stream(q);
    q.set_device();
    hip::check_hip(hipMemcpyAsync(d.raw_storage().values, m.data(), m.size1() * m.size2() * sizeof(T),
                                  hipMemcpyHostToDevice, s.handle()), "hipMemcpyAsync");
    s.synchronize();
    return d;
}

template <class T, class O>
matrix<T, O, cpu_tag> copy_to_cpu(matrix<T, O, hip_tag> const& d) {
    matrix<T, O, cpu_tag> m(d.size1(), d.size2());
    if (m.size1() * m.size2() == 0) return m;
    hip::stream& s = hip::get_stream(d.queue());
    d.queue().set_device();
    hip::detail::download(m.data(), d.raw_storage().values, m.size1() * m.size2() * sizeof(T), s);
    return m;
}

template <class T>
vector<T, hip_tag> copy_to_device(vector<T, cpu_tag> const& v, hip_tag,
                                  hip::device& q = hip::device_manager::instance().default_device()) {
    vector<T, hip_tag> d(v.size(), q);
    if (v.size() == 0) return d;
    hip::stream& s = hip::get_stream(q);
    q.set_device();
    hip::check_hip(hipMemcpyAsync(d.raw_storage().values, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                                  s.handle()), "hipMemcpyAsync");
    s.synchronize();
    return d;
}

template <class E>
vector<typename E::value_type, cpu_tag> copy_to_cpu(vector_expression<E, hip_tag> const& d) {
    typedef typename E::value_type T;
    vector<T, cpu_tag> v(d().size());
    if (v.size() == 0) return v;
    hip::device& q = d().queue();
    auto sd = d().raw_storage();
    vector<T, hip_tag> packed(v.size(), q);  // gathers a strided view
    detail::copy_vector(packed.raw_storage().values, 1, sd.values, sd.stride, v.size(), q);
    hip::stream& s = hip::get_stream(q);
    hip::detail::download(v.data(), packed.raw_storage().values, v.size() * sizeof(T), s);
    return v;
}

}  // namespace remora

#endif

// remora/hip/core.hpp -- tags, errors, devices, streams and buffers of the
// hip_tag front-end, re-provided on the modern HIP runtime API.
//
// Mirrors (same names and meaning, own code):
//   device tags / orientation / triangle / side tags ... include/remora/expression_types.hpp:33-35,
//                                                        include/remora/detail/structure.hpp:50-141
//   hip_exception + check_hip ........................... include/remora/hip/exception.hpp:41-63
//   device / device_manager / stream / get_stream ....... include/remora/hip/device.hpp:41-173
//   buffer<T> ............................................ include/remora/hip/buffer.hpp:41-102
// The cuBLAS handle cache of get_blas (hip/cublas_backend.hpp:442-450) is
// replaced by get_handle(), a per-thread, per-device rmh_handle.
#ifndef REMORA_HIP_CORE_HPP
#define REMORA_HIP_CORE_HPP

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <system_error>
#include <utility>
#include <vector>

#include "../../remora_hip/capi.h"

namespace remora {

// ---------------------------------------------------------------- tags --
struct cpu_tag {};
struct hip_tag {};

struct row_major;
struct column_major;
struct row_major {
    typedef column_major transposed_orientation;
    static constexpr bool is_row_major = true;
};
struct column_major {
    typedef row_major transposed_orientation;
    static constexpr bool is_row_major = false;
};

template <bool Upper, bool Unit>
struct triangular_tag {
    static constexpr bool is_upper = Upper;
    static constexpr bool is_unit = Unit;
    typedef triangular_tag<!Upper, Unit> transposed_orientation;
};
typedef triangular_tag<false, false> lower;
typedef triangular_tag<true, false> upper;
typedef triangular_tag<false, true> unit_lower;
typedef triangular_tag<true, true> unit_upper;

template <bool Left>
struct system_tag {
    static constexpr bool is_left = Left;
    typedef system_tag<!Left> transposed_orientation;
};
typedef system_tag<true> left;
typedef system_tag<false> right;

struct symm_pos_def {};
struct symm_semi_pos_def {};

// -------------------------------------------------------------- errors --
namespace hip {

inline std::error_category const& hip_category() {
    struct cat : std::error_category {
        const char* name() const noexcept override { return "hip"; }
        std::string message(int e) const override { return hipGetErrorString(static_cast<hipError_t>(e)); }
    };
    static cat c;
    return c;
}

inline std::error_category const& rmh_category() {
    struct cat : std::error_category {
        const char* name() const noexcept override { return "remora_hip"; }
        std::string message(int e) const override { return rmh_status_string(static_cast<rmh_status>(e)); }
    };
    static cat c;
    return c;
}

class hip_exception : public std::system_error {
public:
    hip_exception(hipError_t e, std::string const& what) : std::system_error(int(e), hip_category(), what) {}
};

class rmh_exception : public std::system_error {
public:
    rmh_exception(rmh_status s, std::string const& what) : std::system_error(int(s), rmh_category(), what) {}
};

inline void check_hip(hipError_t e, char const* what = "hip call") {
    if (e != hipSuccess) throw hip_exception(e, what);
}

inline void check_rmh(rmh_status s) {
    if (s != RMH_STATUS_SUCCESS) throw rmh_exception(s, rmh_last_error());
}

// ------------------------------------------------------ device / stream --
class device {
public:
    explicit device(int id = 0) : m_id(id) {
        int count = 0;
        check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (id < 0 || id >= count) throw std::invalid_argument("[remora::hip::device] no such device");
    }
    int device_id() const { return m_id; }
    void set_device() const { check_hip(hipSetDevice(m_id), "hipSetDevice"); }
    std::size_t warp_size() const {
        hipDeviceProp_t p;
        check_hip(hipGetDeviceProperties(&p, m_id), "hipGetDeviceProperties");
        return std::size_t(p.warpSize);
    }

private:
    int m_id;
};

// Process-wide device list; device 0 is the default queue of new matrices.
class device_manager {
public:
    static device_manager& instance() {
        static device_manager* m = new device_manager();  // lives until exit
        return *m;
    }
    device& get(int id) { return *m_devices.at(std::size_t(id)); }
    device& default_device() { return get(0); }

private:
    device_manager() {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
        for (int i = 0; i < count; ++i) m_devices.emplace_back(new device(i));
    }
    std::vector<std::unique_ptr<device>> m_devices;
};

// A stream per thread per device (the reference keeps one per thread,
// hip/device.hpp:162-173).  Default priority only.
class stream {
public:
    explicit stream(device const& d) : m_device(d.device_id()) {
        d.set_device();
        check_hip(hipStreamCreate(&m_stream), "hipStreamCreate");
    }
    stream(stream const&) = delete;
    stream& operator=(stream const&) = delete;
    ~stream() {
        if (m_stream) {
            (void)hipSetDevice(m_device);
            (void)hipStreamDestroy(m_stream);
        }
    }
    hipStream_t handle() const { return m_stream; }
    rmh_stream_t abi() const { return reinterpret_cast<rmh_stream_t>(m_stream); }
    void synchronize() const { check_hip(hipStreamSynchronize(m_stream), "hipStreamSynchronize"); }

private:
    int m_device;
    hipStream_t m_stream = nullptr;
};

// The per-thread caches below are never torn down: they live until the
// process ends, so no HIP object is released after the runtime's own
// shutdown has begun.
inline stream& get_stream(device const& d) {
    thread_local std::vector<std::pair<int, stream*>>* streams = new std::vector<std::pair<int, stream*>>();
    for (auto& s : *streams)
        if (s.first == d.device_id()) return *s.second;
    streams->emplace_back(d.device_id(), new stream(d));
    return *streams->back().second;
}

// One rmh_handle per thread per device (replaces get_blas).
inline rmh_handle get_handle(device const& d) {
    thread_local std::vector<std::pair<int, rmh_handle>>* handles = new std::vector<std::pair<int, rmh_handle>>();
    for (auto& o : *handles)
        if (o.first == d.device_id()) return o.second;
    rmh_handle h = nullptr;
    check_rmh(rmh_create(&h, d.device_id()));
    handles->emplace_back(d.device_id(), h);
    return h;
}

// A few device int64 words per thread and device for the status values the
// front end reads back (potrf's info, pstrf's rank, update's failing
// column): allocated on first use and kept until the process ends, so those
// calls allocate nothing after the first.  Calls on one thread use its
// stream in order and read the word back before returning, so one word per
// kind is enough.
enum status_word { status_info = 0, status_rank = 1, status_update = 2, status_words_count = 4 };
inline int64_t* status_words(device const& d, status_word w = status_info) {
    thread_local std::vector<std::pair<int, int64_t*>>* words = new std::vector<std::pair<int, int64_t*>>();
    for (auto& e : *words)
        if (e.first == d.device_id()) return e.second + w;
    d.set_device();
    int64_t* p = nullptr;
    check_hip(hipMalloc(&p, status_words_count * sizeof(int64_t)), "hipMalloc");
    words->emplace_back(d.device_id(), p);
    return p + w;
}
// Wait for the thread's stream and return the word's value.
inline int64_t read_status_word(device const& d, const int64_t* word) {
    stream& s = get_stream(d);
    int64_t h = 0;
    check_hip(hipMemcpyAsync(&h, word, sizeof(h), hipMemcpyDeviceToHost, s.handle()), "hipMemcpyAsync");
    s.synchronize();
    return h;
}

// ------------------------------------------------------------- buffer --
// Owning device allocation (hip/buffer.hpp:41-102).
template <class T>
class buffer {
public:
    buffer() = default;
    buffer(std::size_t size, device& d) : m_device(&d), m_size(size) {
        if (size) {
            d.set_device();
            check_hip(hipMalloc(&m_ptr, size * sizeof(T)), "hipMalloc");
        }
    }
    buffer(buffer const&) = delete;
    buffer& operator=(buffer const&) = delete;
    buffer(buffer&& o) noexcept { swap(o); }
    buffer& operator=(buffer&& o) noexcept {
        buffer tmp(std::move(o));
        swap(tmp);
        return *this;
    }
    ~buffer() {
        if (m_ptr) {
            (void)hipSetDevice(m_device->device_id());
            (void)hipFree(m_ptr);
        }
    }
    void swap(buffer& o) noexcept {
        std::swap(m_device, o.m_device);
        std::swap(m_ptr, o.m_ptr);
        std::swap(m_size, o.m_size);
    }
    T* get() const { return m_ptr; }
    std::size_t size() const { return m_size; }
    device& queue() const { return *m_device; }

private:
    device* m_device = nullptr;
    T* m_ptr = nullptr;
    std::size_t m_size = 0;
};

}  // namespace hip
}  // namespace remora

#endif

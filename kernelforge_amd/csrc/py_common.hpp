// py_common.hpp -- shared glue of the pybind11 modules: NumPy <-> device buffers,
// status-code -> Python exception mapping.  Compute always runs through the C ABI
// of libkf_hip.so (include/kf_hip.h); there is no host fallback.
#pragma once

#include <hip/hip_runtime_api.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kf_hip.h"

namespace py = pybind11;

namespace kfpy {

template <typename T>
using carray = py::array_t<T, py::array::c_style | py::array::forcecast>;

// Map a C-ABI status to the exception type the reference raises.
inline void check(int rc, const char *what) {
    if (rc == KF_OK) return;
    std::string msg = std::string(what) + ": " + kf_last_error();
    switch (rc) {
        case KF_ERR_ARG:
        case KF_ERR_SIGMA:
        case KF_ERR_BASIS:
            throw py::value_error(msg);
        case KF_ERR_NOMEM:
            throw std::bad_alloc();
        default:
            if (rc > 0) throw std::runtime_error(std::string(what) + " failed, info=" + std::to_string(rc));
            throw std::runtime_error(msg);
    }
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// RAII device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) : n(count) {
        if (count) {
            if (hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)) != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
                throw std::bad_alloc();
            }
        }
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void upload(const T *src, size_t count) {
        if (count) hip_check(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(T *dst, size_t count) const {
        if (count) hip_check(hipMemcpy(dst, p, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
};

template <typename T>
DevBuf<T> to_device(const carray<T> &a) {
    DevBuf<T> d((size_t)a.size());
    d.upload(a.data(), (size_t)a.size());
    return d;
}

template <typename T>
DevBuf<T> to_device(const std::vector<T> &v) {
    DevBuf<T> d(v.size());
    d.upload(v.data(), v.size());
    return d;
}

inline int clamp_atoms(int n, int ma) { return n < 0 ? 0 : (n > ma ? ma : n); }

inline long naq_of(const carray<int> &N, int ma) {
    long s = 0;
    auto r = N.unchecked<1>();
    for (py::ssize_t i = 0; i < r.shape(0); ++i) s += 3 * clamp_atoms(r(i), ma);
    return s;
}

}  // namespace kfpy

// df::DeviceField1D / df::DeviceField2D -- header-only C++ facade over the C ABI.
//
// Reference interface: include/utils/device_fields.cuh:303-612 (thrust-backed SoA device
// arrays with from_host/to_host/fill/scale/copy_from/ptr/xptr/yptr).  Here a field is either
//   * owned: device memory from dpmd_malloc (per-system parameters such as dt, flags), or
//   * a view of a context field (pos, vel, rad, box_size, ...), looked up through the
//     context on every access so that reallocation by the library never leaves it stale.
// No HIP header is needed by callers; every device operation goes through dpmd.h.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "dpmd.h"

namespace df {

namespace detail {
inline void check(int status, const dpmd_ctx* ctx = nullptr) {
    if (status != DPMD_OK) {
        const char* m = dpmd_last_error(ctx);
        const std::string msg = m ? m : "dpmd error";
        if (status == DPMD_ERR_ARGUMENT) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
}
}  // namespace detail

template <class T>
class DeviceField1D {
    static_assert(std::is_same<T, double>::value || std::is_same<T, int>::value ||
                      std::is_same<T, unsigned int>::value || std::is_same<T, unsigned char>::value,
                  "DeviceField1D supports double, int, unsigned int and unsigned char");

public:
    DeviceField1D() = default;
    explicit DeviceField1D(int n) { resize(n); }
    // view of a context field
    DeviceField1D(dpmd_ctx* ctx, int field) : ctx_(ctx), field_(field) {}
    DeviceField1D(const DeviceField1D& o) { *this = o; }
    DeviceField1D& operator=(const DeviceField1D& o) {
        if (this == &o) return *this;
        if (o.is_view()) { release(); ctx_ = o.ctx_; field_ = o.field_; return *this; }
        ctx_ = nullptr; field_ = -1;
        resize(o.size());
        copy_from(o);
        return *this;
    }
    DeviceField1D(DeviceField1D&& o) noexcept { swap_with(o); }
    DeviceField1D& operator=(DeviceField1D&& o) noexcept { swap_with(o); return *this; }
    ~DeviceField1D() { release(); }

    bool is_view() const { return ctx_ != nullptr; }

    int size() const {
        if (!is_view()) return n_;
        int64_t n = 0;
        detail::check(dpmd_field(ctx_, field_, nullptr, &n), ctx_);
        return (int)n;
    }

    T* ptr() const {
        if (!is_view()) return p_;
        void* p = nullptr;
        detail::check(dpmd_field(ctx_, field_, &p, nullptr), ctx_);
        return (T*)p;
    }

    void resize(int n) {
        if (is_view()) {
            if (n != size()) throw std::runtime_error("DeviceField1D::resize: context fields are sized by allocate_*");
            return;
        }
        if (n == n_) return;
        release();
        void* p = nullptr;
        detail::check(dpmd_malloc(&p, (int64_t)(n > 0 ? n : 1) * (int64_t)sizeof(T)));
        p_ = (T*)p;
        n_ = n;
    }

    void from_host(const std::vector<T>& h) {
        if (!is_view()) resize((int)h.size());
        if ((int)h.size() != size()) throw std::invalid_argument("DeviceField1D::from_host: size mismatch");
        if (!h.empty()) detail::check(dpmd_memcpy(ptr(), h.data(), (int64_t)h.size() * sizeof(T), 0));
    }

    void to_host(std::vector<T>& h) const {
        h.resize(size());
        if (!h.empty()) detail::check(dpmd_memcpy(h.data(), ptr(), (int64_t)h.size() * sizeof(T), 1));
    }

    std::vector<T> to_host() const { std::vector<T> h; to_host(h); return h; }

    void fill(T v) {
        const int n = size();
        if (n == 0) return;
        if (std::is_same<T, double>::value) detail::check(dpmd_fill_f64(ptr(), n, (double)v));
        else if (sizeof(T) == 4) detail::check(dpmd_fill_i32(ptr(), n, (int32_t)v));
        else { std::vector<T> h(n, v); from_host(h); }
    }

    void scale(T s) {   // k_scale (device_fields.cuh:34)
        std::vector<T> h;
        to_host(h);
        for (auto& v : h) v = (T)(v * s);
        from_host(h);
    }

    void copy_from(const DeviceField1D& o) {
        if (o.size() != size()) throw std::invalid_argument("DeviceField1D::copy_from: size mismatch");
        if (size()) detail::check(dpmd_memcpy(ptr(), o.ptr(), (int64_t)size() * sizeof(T), 2));
    }

    T get_element(int i) const {
        T v{};
        detail::check(dpmd_memcpy(&v, ptr() + i, sizeof(T), 1));
        return v;
    }
    void set_element(int i, T v) { detail::check(dpmd_memcpy(ptr() + i, &v, sizeof(T), 0)); }

private:
    void release() {
        if (p_) (void)dpmd_free(p_);
        p_ = nullptr; n_ = 0;
    }
    void swap_with(DeviceField1D& o) noexcept {
        std::swap(p_, o.p_); std::swap(n_, o.n_); std::swap(ctx_, o.ctx_); std::swap(field_, o.field_);
    }

    T* p_ = nullptr;
    int n_ = 0;
    dpmd_ctx* ctx_ = nullptr;
    int field_ = -1;
};

// Separate SoA x and y arrays (device_fields.cuh:468-612).
template <class T>
class DeviceField2D {
public:
    DeviceField1D<T> x, y;
    DeviceField2D() = default;
    explicit DeviceField2D(int n) : x(n), y(n) {}
    DeviceField2D(dpmd_ctx* ctx, int fx, int fy) : x(ctx, fx), y(ctx, fy) {}
    int size() const { return x.size(); }
    void resize(int n) { x.resize(n); y.resize(n); }
    T* xptr() const { return x.ptr(); }
    T* yptr() const { return y.ptr(); }
    void from_host(const std::vector<T>& hx, const std::vector<T>& hy) { x.from_host(hx); y.from_host(hy); }
    void to_host(std::vector<T>& hx, std::vector<T>& hy) const { x.to_host(hx); y.to_host(hy); }
    void fill(T vx, T vy) { x.fill(vx); y.fill(vy); }
    void copy_from(const DeviceField2D& o) { x.copy_from(o.x); y.copy_from(o.y); }
};

}  // namespace df

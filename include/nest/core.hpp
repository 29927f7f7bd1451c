// nest core types — index spaces, SoA fields, execution tags.
//
// API-compatible with the reference's include/nest/core.hpp (same names,
// signatures and semantics; own implementation). The layout contract the
// GPU backend relies on: field_t stores variable-major, and within a
// variable the LAST index is contiguous (row-major), exactly as the reference
// (core.hpp:206-221, :310-392), so a patch array uploads to the device as is.
#pragma once
#include <algorithm>
#include <array>
#include <cstddef>
#include <cstring>
#include <type_traits>
#include <utility>

#if defined(_OPENMP)
#include <omp.h>
#endif

namespace nest {

enum class field_init_t { zeroed, uninitialized };

// ---------------------------------------------------------------------------
// vec_t — fixed-size arithmetic vector (reference core.hpp:22-87)
// ---------------------------------------------------------------------------
template <typename T, std::size_t N>
struct vec_t {
    std::array<T, N> _data{};

    constexpr vec_t() = default;
    template <typename... Args>
        requires(sizeof...(Args) == N && (std::is_convertible_v<Args, T> && ...))
    constexpr vec_t(Args... args) : _data{static_cast<T>(args)...} {}

    constexpr T& operator[](std::size_t i) { return _data[i]; }
    constexpr const T& operator[](std::size_t i) const { return _data[i]; }
    constexpr std::size_t size() const { return N; }
    constexpr auto begin() { return _data.begin(); }
    constexpr auto end() { return _data.end(); }
    constexpr auto begin() const { return _data.begin(); }
    constexpr auto end() const { return _data.end(); }
    constexpr bool operator==(const vec_t&) const = default;
};

template <std::size_t N> using ivec_t = vec_t<int, N>;
template <std::size_t N> using uvec_t = vec_t<unsigned int, N>;
template <std::size_t N> using dvec_t = vec_t<double, N>;

template <typename... Args>
constexpr ivec_t<sizeof...(Args)> ivec(Args... a) { return ivec_t<sizeof...(Args)>{static_cast<int>(a)...}; }
template <typename... Args>
constexpr uvec_t<sizeof...(Args)> uvec(Args... a) {
    return uvec_t<sizeof...(Args)>{static_cast<unsigned int>(a)...};
}

template <typename T, std::size_t N>
constexpr vec_t<T, N> operator+(const vec_t<T, N>& a, const vec_t<T, N>& b) {
    vec_t<T, N> r;
    for (std::size_t i = 0; i < N; ++i) r[i] = a[i] + b[i];
    return r;
}
template <typename T, std::size_t N>
constexpr vec_t<T, N> operator-(const vec_t<T, N>& a, const vec_t<T, N>& b) {
    vec_t<T, N> r;
    for (std::size_t i = 0; i < N; ++i) r[i] = a[i] - b[i];
    return r;
}
template <typename T, std::size_t N, typename S>
constexpr vec_t<T, N> operator*(const vec_t<T, N>& v, S s) {
    vec_t<T, N> r;
    for (std::size_t i = 0; i < N; ++i) r[i] = v[i] * static_cast<T>(s);
    return r;
}
template <typename T, std::size_t N>
constexpr T product(const vec_t<T, N>& v) {
    T p{1};
    for (std::size_t i = 0; i < N; ++i) p *= v[i];
    return p;
}

// ---------------------------------------------------------------------------
// index_space_t — box [start, start + shape), iterated row-major (last index
// fastest), reference core.hpp:93-199
// ---------------------------------------------------------------------------
template <std::size_t Dim>
struct index_space_t {
    ivec_t<Dim> _start{};
    uvec_t<Dim> _shape{};

    constexpr index_space_t() = default;
    constexpr index_space_t(ivec_t<Dim> s, uvec_t<Dim> n) : _start(s), _shape(n) {}

    struct iterator {
        ivec_t<Dim> _current, _start;
        uvec_t<Dim> _shape;
        constexpr ivec_t<Dim> operator*() const { return _current; }
        constexpr iterator& operator++() {
            for (int d = static_cast<int>(Dim) - 1; d >= 0; --d) {
                if (++_current[d] < _start[d] + static_cast<int>(_shape[d])) return *this;
                _current[d] = _start[d];
            }
            _current[0] = _start[0] + static_cast<int>(_shape[0]);  // past the end
            return *this;
        }
        constexpr bool operator==(const iterator& o) const { return _current == o._current; }
        constexpr bool operator!=(const iterator& o) const { return !(*this == o); }
    };

    constexpr iterator begin() const {
        // an empty box starts at its end
        for (std::size_t d = 0; d < Dim; ++d)
            if (_shape[d] == 0) return end();
        return iterator{_start, _start, _shape};
    }
    constexpr iterator end() const {
        ivec_t<Dim> e = _start;
        e[0] = _start[0] + static_cast<int>(_shape[0]);
        return iterator{e, _start, _shape};
    }
};

template <std::size_t Dim> constexpr ivec_t<Dim> start(const index_space_t<Dim>& s) { return s._start; }
template <std::size_t Dim> constexpr uvec_t<Dim> shape(const index_space_t<Dim>& s) { return s._shape; }
template <std::size_t Dim> constexpr std::size_t size(const index_space_t<Dim>& s) {
    return static_cast<std::size_t>(product(s._shape));
}
template <std::size_t Dim> constexpr ivec_t<Dim> upper(const index_space_t<Dim>& s) {
    ivec_t<Dim> u;
    for (std::size_t d = 0; d < Dim; ++d) u[d] = s._start[d] + static_cast<int>(s._shape[d]);
    return u;
}
template <std::size_t Dim>
constexpr index_space_t<Dim> index_space(ivec_t<Dim> s, uvec_t<Dim> n) { return {s, n}; }

template <std::size_t Dim>
constexpr index_space_t<Dim> expand(const index_space_t<Dim>& s, int amount) {
    index_space_t<Dim> r = s;
    for (std::size_t d = 0; d < Dim; ++d) {
        r._start[d] -= amount;
        r._shape[d] += static_cast<unsigned int>(2 * amount);
    }
    return r;
}

// Part part_idx of n_parts along `axis`: the first (total % n) parts get one
// extra cell (reference core.hpp:184-199).
template <std::size_t Dim>
constexpr index_space_t<Dim> subspace(const index_space_t<Dim>& s, int n_parts, int part_idx, int axis) {
    const int total = static_cast<int>(s._shape[axis]);
    const int base = total / n_parts, extra = total % n_parts;
    index_space_t<Dim> r = s;
    r._start[axis] += part_idx * base + std::min(part_idx, extra);
    r._shape[axis] = static_cast<unsigned int>(base + (part_idx < extra ? 1 : 0));
    return r;
}

// Row-major offset of idx within the box (reference core.hpp:206-215)
template <std::size_t Dim>
constexpr std::size_t linear_index(const index_space_t<Dim>& s, ivec_t<Dim> idx) {
    std::size_t off = 0;
    for (std::size_t d = 0; d < Dim; ++d) off = off * s._shape[d] + static_cast<std::size_t>(idx[d] - s._start[d]);
    return off;
}
template <std::size_t Dim>
constexpr std::size_t soa_offset(const index_space_t<Dim>& s, int var, ivec_t<Dim> idx) {
    return static_cast<std::size_t>(var) * size(s) + linear_index(s, idx);
}

// ---------------------------------------------------------------------------
// md_view_t — pointer + box + element strides (reference core.hpp:227-304)
// ---------------------------------------------------------------------------
template <typename T, std::size_t Dim>
struct md_view_t {
    T* _data = nullptr;
    index_space_t<Dim> _space;
    uvec_t<Dim> _strides;

    md_view_t() = default;
    md_view_t(T* data, index_space_t<Dim> sp) : _data(data), _space(sp) {
        unsigned int st = 1;
        for (int d = static_cast<int>(Dim) - 1; d >= 0; --d) {
            _strides[d] = st;
            st *= sp._shape[d];
        }
    }
    md_view_t(T* data, index_space_t<Dim> sp, uvec_t<Dim> st) : _data(data), _space(sp), _strides(st) {}

    std::size_t offset_of(ivec_t<Dim> idx) const {
        std::size_t off = 0;
        for (std::size_t d = 0; d < Dim; ++d) off += static_cast<std::size_t>(idx[d] - _space._start[d]) * _strides[d];
        return off;
    }
    T& operator()(ivec_t<Dim> idx) { return _data[offset_of(idx)]; }
    const T& operator()(ivec_t<Dim> idx) const { return _data[offset_of(idx)]; }
    T* data_at(ivec_t<Dim> idx) { return _data + offset_of(idx); }
    md_view_t subview(index_space_t<Dim> sub) const { return {_data + offset_of(sub._start), sub, _strides}; }
};

template <typename T, std::size_t Dim> index_space_t<Dim> space(const md_view_t<T, Dim>& v) { return v._space; }
template <typename T, std::size_t Dim> uvec_t<Dim> strides(const md_view_t<T, Dim>& v) { return v._strides; }
template <typename T, std::size_t Dim>
bool is_contiguous(const md_view_t<T, Dim>& v) {
    std::size_t expect = 1;
    for (int d = static_cast<int>(Dim) - 1; d >= 0; --d) {
        if (v._strides[d] != expect) return false;
        expect *= v._space._shape[d];
    }
    return true;
}

// ---------------------------------------------------------------------------
// field_t — nvars variables over one box, SoA (reference core.hpp:310-392)
// ---------------------------------------------------------------------------
template <typename T, std::size_t Dim>
struct field_t {
    T* _data = nullptr;
    index_space_t<Dim> _space;
    std::size_t _nvars = 0;
    std::size_t _stride = 0;
    bool _owns_data = false;

    field_t() = default;
    field_t(std::size_t nvars, index_space_t<Dim> sp, field_init_t init = field_init_t::zeroed)
        : _space(sp), _nvars(nvars), _stride(size(sp)), _owns_data(true) {
        _data = init == field_init_t::zeroed ? new T[nvars * _stride]() : new T[nvars * _stride];
    }
    field_t(T* data, std::size_t nvars, index_space_t<Dim> sp)
        : _data(data), _space(sp), _nvars(nvars), _stride(size(sp)), _owns_data(false) {}
    ~field_t() { release(); }
    field_t(field_t&& o) noexcept { steal(o); }
    field_t& operator=(field_t&& o) noexcept {
        if (this != &o) {
            release();
            steal(o);
        }
        return *this;
    }
    field_t(const field_t&) = delete;
    field_t& operator=(const field_t&) = delete;

    md_view_t<T, Dim> operator[](std::size_t v) { return {_data + v * _stride, _space}; }
    md_view_t<const T, Dim> operator[](std::size_t v) const { return {_data + v * _stride, _space}; }
    T& operator()(std::size_t v, ivec_t<Dim> idx) { return _data[v * _stride + linear_index(_space, idx)]; }
    const T& operator()(std::size_t v, ivec_t<Dim> idx) const { return _data[v * _stride + linear_index(_space, idx)]; }
    T* var_data(std::size_t v) { return _data + v * _stride; }
    const T* var_data(std::size_t v) const { return _data + v * _stride; }
    std::size_t nvars() const { return _nvars; }
    std::size_t stride() const { return _stride; }

  private:
    void release() {
        if (_owns_data) delete[] _data;
        _data = nullptr;
    }
    void steal(field_t& o) {
        _data = o._data;
        _space = o._space;
        _nvars = o._nvars;
        _stride = o._stride;
        _owns_data = o._owns_data;
        o._data = nullptr;
        o._owns_data = false;
    }
};

// state_t — conserved variables of one patch (reference core.hpp:398-426)
template <typename T, std::size_t Dim>
struct state_t {
    field_t<T, Dim> conserved;
    double time = 0.0;

    state_t() = default;
    state_t(std::size_t nvars, index_space_t<Dim> sp) : conserved(nvars, sp) {}
    md_view_t<T, Dim> operator[](std::size_t v) { return conserved[v]; }
    md_view_t<const T, Dim> operator[](std::size_t v) const { return conserved[v]; }
    T& operator()(std::size_t v, ivec_t<Dim> idx) { return conserved(v, idx); }
    const T& operator()(std::size_t v, ivec_t<Dim> idx) const { return conserved(v, idx); }
    std::size_t nvars() const { return conserved.nvars(); }
};

// ---------------------------------------------------------------------------
// execution policies (reference core.hpp:432-440). exec::hip (the MI355X
// backend) is declared in backend_hip.hpp; cuda_t is kept for source
// compatibility and has no executor, as in the reference.
// ---------------------------------------------------------------------------
namespace exec {
struct cpu_t {};
struct omp_t {};
struct cuda_t {};
inline constexpr cpu_t cpu;
inline constexpr omp_t omp;
inline constexpr cuda_t cuda;
}  // namespace exec

template <std::size_t Dim, typename F>
void for_each(const index_space_t<Dim>& s, F&& f, exec::cpu_t) {
    for (auto idx : s) f(idx);
}

// OpenMP over the flattened box (serial without OpenMP), reference core.hpp:456-477
template <std::size_t Dim, typename F>
void for_each(const index_space_t<Dim>& s, F&& f, exec::omp_t) {
    const long n = static_cast<long>(size(s));
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        ivec_t<Dim> idx;
        std::size_t rem = static_cast<std::size_t>(k);
        for (int d = static_cast<int>(Dim) - 1; d >= 0; --d) {
            idx[d] = s._start[d] + static_cast<int>(rem % s._shape[d]);
            rem /= s._shape[d];
        }
        f(idx);
    }
}

template <std::size_t Dim, typename F>
void for_each(const index_space_t<Dim>& s, F&& f) {
    for_each(s, std::forward<F>(f), exec::cpu);
}

}  // namespace nest

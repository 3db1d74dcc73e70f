// user_rule.hpp -- user rules and observables compiled as device functors.
//
// The reference's extension points are a subclass of quids::rule with the
// virtual methods get_num_child / populate_child / populate_child_simple /
// hasher (src/quids.hpp:105-146, README.md:112-180) and any observable_t
// callable (src/quids.hpp:83, 208-234).  A virtual method cannot run on the
// GPU, so a user rule here is a trivially copyable struct R whose methods
// have the reference's signatures and are __host__ __device__ (QD_HD):
//
//   struct my_rule {
//       QD_HD void get_num_child(const char *parent_begin, const char *parent_end,
//                                uint32_t &num_child, uint32_t &max_child_size) const;
//       QD_HD void populate_child(const char *parent_begin, const char *parent_end,
//                                 char *child_begin, uint32_t child_id, uint32_t &size,
//                                 quids::dmag &mag) const;
//       // optional (defaults as in the reference):
//       QD_HD void populate_child_simple(const char *parent_begin, const char *parent_end,
//                                        char *child_begin, uint32_t child_id) const;
//       QD_HD uint64_t hasher(const char *object_begin, const char *object_end) const;
//   };
//
// quids::device_rule<R> wraps it as a quids::rule: quids::simulate(it, &rule,
// next, sy) launches R's kernels (below, compiled into the caller's program
// by hipcc --offload-arch=gfx950) for the rule-specific phases and the
// library's for the rest (qd_simulate_user).  The default hasher is the
// reference's std::hash<std::string_view> (MurmurHash64A), as for the
// built-in rules, so a user restatement of a built-in rule reproduces it bit
// for bit.  populate_child must write at most the max_child_size bytes that
// get_num_child reported for the parent.
//
// An observable is a functor F with QD_HD double operator()(const char
// *object_begin, const char *object_end) const; quids::device_observable<F>
// evaluates it per object on the device for iteration::average_value.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "csrc/qd_rules.hpp"
#include "quids.hpp"

namespace quids {

/// The magnitude a device populate_child multiplies (the reference's mag_t,
/// std::complex<double>; products in the order (a.re b.re - a.im b.im,
/// a.re b.im + a.im b.re), as the reference's rules and the built-in ones).
struct dmag {
    double re, im;
    QD_HD dmag &operator*=(double s) {
        re *= s;
        im *= s;
        return *this;
    }
    QD_HD dmag &operator*=(dmag b) {
        const double r = re * b.re - im * b.im, i = re * b.im + im * b.re;
        re = r;
        im = i;
        return *this;
    }
    QD_HD dmag &operator+=(dmag b) {
        re += b.re;
        im += b.im;
        return *this;
    }
};
QD_HD dmag operator*(dmag a, dmag b) { return a *= b; }
QD_HD dmag operator*(dmag a, double s) { return a *= s; }
QD_HD dmag operator*(double s, dmag a) { return a *= s; }
QD_HD dmag operator+(dmag a, dmag b) { return a += b; }

namespace user_detail {

template <class R, class = void>
struct has_simple : std::false_type {};
template <class R>
struct has_simple<R, std::void_t<decltype(std::declval<const R &>().populate_child_simple(
                         (const char *)0, (const char *)0, (char *)0, 0u))>> : std::true_type {};
template <class R, class = void>
struct has_hasher : std::false_type {};
template <class R>
struct has_hasher<R, std::void_t<decltype(std::declval<const R &>().hasher((const char *)0, (const char *)0))>>
    : std::true_type {};

template <class R>
QD_HD uint64_t hash_of(const R &r, const char *b, const char *e) {
    if constexpr (has_hasher<R>::value) return uint64_t(r.hasher(b, e));
    else return qd::hash_bytes(reinterpret_cast<const uint8_t *>(b), uint32_t(e - b));
}

template <class R>
QD_HD void populate_simple(const R &r, const char *pb, const char *pe, char *child, uint32_t id) {
    if constexpr (has_simple<R>::value) {
        r.populate_child_simple(pb, pe, child, id);
    } else {
        uint32_t size = 0;
        dmag m{1.0, 0.0};
        r.populate_child(pb, pe, child, id, size, m);
    }
}

// largest p < n with cb[p] <= c (cb: n + 1 ascending offsets, cb[n] > c)
__device__ inline uint64_t parent_of(const uint64_t *cb, uint64_t n, uint64_t c) {
    uint64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (cb[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

inline unsigned grid_of(uint64_t n) { return unsigned(std::min<uint64_t>((n + 255) / 256, 65536)); }

// src/quids.hpp:557-567: num_child and max_child_size of every parent
template <class R>
__global__ __launch_bounds__(256) void k_num_child(R r, qd_objects_view P, uint32_t *nc, uint32_t *ms) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += uint64_t(gridDim.x) * blockDim.x) {
        const char *b = reinterpret_cast<const char *>(P.objects) + P.begin[i];
        uint32_t n = 0, m = 0;
        r.get_num_child(b, b + P.size[i], n, m);
        nc[i] = n;
        ms[i] = m;
    }
}

// src/quids.hpp:705-719: each thread builds its candidates' bytes in its own
// slice of scratch, then hashes them
template <class R>
__global__ __launch_bounds__(256) void k_generate(R r, qd_objects_view P, const uint64_t *cb, uint64_t C,
                                                  uint8_t *scratch, uint64_t slice, uint64_t *hash, double *mag,
                                                  uint32_t *size) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    char *child = reinterpret_cast<char *>(scratch + t * slice);
    for (uint64_t c = t; c < C; c += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t p = parent_of(cb, P.n, c);
        const char *b = reinterpret_cast<const char *>(P.objects) + P.begin[p];
        dmag m{P.mag[2 * p], P.mag[2 * p + 1]};
        uint32_t sz = 0;
        r.populate_child(b, b + P.size[p], child, uint32_t(c - cb[p]), sz, m);
        hash[c] = hash_of(r, child, child + sz);
        mag[2 * c] = m.re;
        mag[2 * c + 1] = m.im;
        size[c] = sz;
    }
}

// src/quids.hpp:956-967: the survivors' bytes
template <class R>
__global__ __launch_bounds__(256) void k_populate(R r, qd_objects_view P, const uint64_t *cb, const uint64_t *cand,
                                                  uint64_t n, const uint64_t *out_begin, uint8_t *out) {
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t c = cand[k];
        const uint64_t p = parent_of(cb, P.n, c);
        const char *b = reinterpret_cast<const char *>(P.objects) + P.begin[p];
        populate_simple(r, b, b + P.size[p], reinterpret_cast<char *>(out + out_begin[k]), uint32_t(c - cb[p]));
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_hash_objects(R r, qd_objects_view P, uint64_t *h) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += uint64_t(gridDim.x) * blockDim.x) {
        const char *b = reinterpret_cast<const char *>(P.objects) + P.begin[i];
        h[i] = hash_of(r, b, b + P.size[i]);
    }
}

template <class F>
__global__ __launch_bounds__(256) void k_observable(F f, qd_objects_view P, double *obs) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += uint64_t(gridDim.x) * blockDim.x) {
        const char *b = reinterpret_cast<const char *>(P.objects) + P.begin[i];
        obs[i] = f(b, b + P.size[i]);
    }
}

inline int launched() { return hipGetLastError() == hipSuccess ? 0 : 1; }

// the callbacks of qd_user_rule for R (self: a const R *)
template <class R>
struct callbacks {
    static int num_child(void *self, const qd_objects_view *P, uint32_t *nc, uint32_t *ms, void *stream) {
        if (P->n == 0) return 0;
        hipLaunchKernelGGL(k_num_child<R>, dim3(grid_of(P->n)), dim3(256), 0, hipStream_t(stream),
                           *static_cast<const R *>(self), *P, nc, ms);
        return launched();
    }
    static int generate(void *self, const qd_objects_view *P, const uint64_t *cb, uint64_t C, uint32_t max_child,
                        uint8_t *scratch, uint64_t scratch_bytes, uint64_t *hash, double *mag, uint32_t *size,
                        void *stream) {
        if (C == 0) return 0;
        const uint64_t slice = (uint64_t(max_child ? max_child : 1) + 15) / 16 * 16;
        const uint64_t threads = scratch_bytes / slice / 256 * 256;
        if (threads == 0) return 2;
        hipLaunchKernelGGL(k_generate<R>, dim3(unsigned(threads / 256)), dim3(256), 0, hipStream_t(stream),
                           *static_cast<const R *>(self), *P, cb, C, scratch, slice, hash, mag, size);
        return launched();
    }
    static int populate(void *self, const qd_objects_view *P, const uint64_t *cb, const uint64_t *cand, uint64_t n,
                        const uint64_t *out_begin, uint8_t *out, void *stream) {
        if (n == 0) return 0;
        hipLaunchKernelGGL(k_populate<R>, dim3(grid_of(n)), dim3(256), 0, hipStream_t(stream),
                           *static_cast<const R *>(self), *P, cb, cand, n, out_begin, out);
        return launched();
    }
    static int hash_objects(void *self, const qd_objects_view *P, uint64_t *h, void *stream) {
        if (P->n == 0) return 0;
        hipLaunchKernelGGL(k_hash_objects<R>, dim3(grid_of(P->n)), dim3(256), 0, hipStream_t(stream),
                           *static_cast<const R *>(self), *P, h);
        return launched();
    }
};

}  // namespace user_detail

/// The C-ABI table of a user rule (r must outlive its use).
template <class R>
qd_user_rule make_user_rule(const R *r) {
    static_assert(std::is_trivially_copyable<R>::value, "a device rule is passed to its kernels by value");
    qd_user_rule u;
    u.self = const_cast<void *>(static_cast<const void *>(r));
    u.num_child = &user_detail::callbacks<R>::num_child;
    u.generate = &user_detail::callbacks<R>::generate;
    u.populate = &user_detail::callbacks<R>::populate;
    u.hash_objects = &user_detail::callbacks<R>::hash_objects;
    return u;
}

/// A user rule as a quids::rule: simulated on the GPU through R's kernels;
/// the virtual methods run R's methods on the host.
template <class R>
class device_rule : public rule {
    R r_;

public:
    explicit device_rule(const R &r) : r_(r) {}
    const R &functor() const { return r_; }
    void get_num_child(char const *b, char const *e, uint &num_child, uint &max_child_size) const override {
        uint32_t n = 0, m = 0;
        r_.get_num_child(b, e, n, m);
        num_child = n;
        max_child_size = m;
    }
    void populate_child(char const *b, char const *e, char *const child, uint const child_id, uint &size,
                        mag_t &mag) const override {
        dmag m{mag.real(), mag.imag()};
        uint32_t sz = 0;
        r_.populate_child(b, e, child, child_id, sz, m);
        size = sz;
        mag = mag_t(m.re, m.im);
    }
    void populate_child_simple(char const *b, char const *e, char *const child, uint const child_id) const override {
        user_detail::populate_simple(r_, b, e, child, child_id);
    }
    size_t hasher(char const *b, char const *e) const override { return size_t(user_detail::hash_of(r_, b, e)); }
    bool user_rule(qd_user_rule &u) const override {
        u = make_user_rule(&r_);
        return true;
    }
};

/// Fills obs (device, one double per object) with f of every object of s on
/// the context's stream.
template <class F>
int eval_observable(qd_ctx *ctx, qd_state *s, const F &f, double *obs) {
    uint64_t n = 0, bytes = 0;
    double t = 0;
    void *objects = nullptr, *stream = nullptr;
    uint64_t *begin = nullptr;
    uint32_t *size = nullptr;
    double *mag = nullptr;
    int rc = qd_state_info(s, &n, &bytes, &t);
    if (rc == QD_OK) rc = qd_state_device_ptrs(s, &objects, &begin, &size, &mag);
    if (rc == QD_OK) rc = qd_ctx_stream(ctx, &stream);
    if (rc != QD_OK || n == 0) return rc;
    const qd_objects_view P{static_cast<const uint8_t *>(objects), begin, size, mag, n};
    hipLaunchKernelGGL(user_detail::k_observable<F>, dim3(user_detail::grid_of(n)), dim3(256), 0,
                       hipStream_t(stream), f, P, obs);
    return user_detail::launched() ? QD_ERR_HIP : QD_OK;
}

/// An observable_t evaluated on the device (iteration::average_value,
/// src/quids.hpp:208-234): sum f(object) |m|^2 in the library's fixed
/// summation order.
template <class F>
class device_observable : public device_observable_base {
    F f_;

public:
    explicit device_observable(const F &f) : f_(f) {}
    double average(qd_ctx *ctx, qd_state *s) const override {
        uint64_t n = 0, bytes = 0;
        double t = 0;
        detail::check(qd_state_info(s, &n, &bytes, &t), "qd_state_info");
        double *obs = nullptr;
        if (n && hipMalloc(&obs, n * sizeof(double)) != hipSuccess) throw std::runtime_error("average_value: hipMalloc");
        double v = 0;
        int rc = eval_observable(ctx, s, f_, obs);
        if (rc == QD_OK) rc = qd_average_value_of(ctx, s, obs, &v, nullptr);
        if (obs) (void)hipFree(obs);
        detail::check(rc, "average_value");
        return v;
    }
};

}  // namespace quids

// quids.hpp -- drop-in C++ API of the QuIDS rule-application step on MI355X.
//
// Mirrors the reference's header-only API (src/quids.hpp): the globals
// (59-75), quids::rule / rule_t (105-146), iteration / it_t (149-335),
// symbolic_iteration / sy_it_t (338-429), modifier_t (86) and the two
// simulate overloads (436-543).  Every step runs on the GPU through the C ABI
// of include/quids_hip.h (link with -lquids_hip); there is no host fallback:
// simulate() with a rule or modifier that has no device implementation
// throws std::runtime_error.  Errors of the library are thrown as
// std::runtime_error, as the reference throws on allocation failure
// (src/utils/vector.hpp:124-127).
//
// Differences a user of the reference may notice:
//  * results follow the canonical choices of the library (representative =
//    smallest candidate index, pairwise merge order, ties at a truncation
//    cut broken by hash, next state in ascending representative order);
//  * with simple_truncation false, the truncation keeps the smallest
//    u / |m|^2 as the reference does, but u is a Philox draw keyed by the
//    object's hash and truncation_seed, so runs are reproducible;
//  * it_t objects live in device memory; reading them (get_object,
//    average_value) downloads a host copy.
#pragma once

#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "../include/quids_hip.h"

namespace quids {

typedef unsigned uint;

/// amount of bytes objects are aligned to (src/quids.hpp:60)
inline uint align_byte_length = 8;
/// objects with a smaller probability are dropped (src/quids.hpp:62)
inline double tolerance = 1e-30;
/// top-|m|^2 truncation (src/quids.hpp:69-75); false: probabilistic
inline bool simple_truncation = true;
/// key of the probabilistic truncation's draws (no reference counterpart)
inline uint64_t truncation_seed = 0;
/// memory safety margin of max_num_object == 0 (src/quids.hpp:64)
inline float safety_margin = 0.2f;
/// memory equalization factor of the next state (src/quids.hpp:66)
inline float equalize_factor = 0.25f;
/// CPU load-balancing granularity (src/quids.hpp:68); read by the QCGD
/// command-line flags, no effect on the device step
inline int load_balancing_bucket_per_thread = 32;

typedef std::complex<double> mag_t;
typedef class iteration it_t;
typedef class symbolic_iteration sy_it_t;
typedef class rule rule_t;
typedef std::function<double(char const *object_begin, char const *object_end)> observable_t;
typedef std::function<void(const char *step)> debug_t;

namespace detail {
inline void check(int rc, const char *where) {
    if (rc != QD_OK) throw std::runtime_error(std::string(where) + ": " + qd_status_string(rc) + ": " + qd_last_error());
}
/// The process-wide device context (device 0), created on first use.
inline qd_ctx *context() {
    struct Holder {
        qd_ctx *c = nullptr;
        Holder() { check(qd_ctx_create(&c, 0), "qd_ctx_create"); }
        ~Holder() { qd_ctx_destroy(c); }
    };
    static Holder h;
    return h.c;
}
inline qd_params params() {
    qd_params p;
    p.tolerance = tolerance;
    p.align_byte_length = align_byte_length;
    p.simple_truncation = simple_truncation ? 1 : 0;
    p.reserved = 0;
    p.seed = truncation_seed;
    p.safety_margin = safety_margin;
    p.equalize_factor = equalize_factor;
    return p;
}
}  // namespace detail

/// utility: alignment padding after an object of `size` bytes (src/quids.hpp:93-102)
inline uint get_alignment_offset(const uint size) {
    if (align_byte_length <= 1) return 0;
    uint off = align_byte_length - size % align_byte_length;
    return off == align_byte_length ? 0 : off;
}

/// A dynamic (rule), src/quids.hpp:105-146.  Rules with a device
/// implementation override device_rule(); user rules that only implement the
/// host methods cannot be simulated on the GPU.
class rule {
public:
    rule() {}
    virtual ~rule() {}
    virtual void get_num_child(char const *parent_begin, char const *parent_end, uint &num_child,
                               uint &max_child_size) const = 0;
    virtual void populate_child(char const *parent_begin, char const *parent_end, char *const child_begin,
                                uint const child_id, uint &size, mag_t &mag) const = 0;
    virtual void populate_child_simple(char const *parent_begin, char const *parent_end, char *const child_begin,
                                       uint const child_id) const {
        uint size;
        mag_t mag;
        populate_child(parent_begin, parent_end, child_begin, child_id, size, mag);
    }
    virtual size_t hasher(char const *object_begin, char const *object_end) const {
        return std::hash<std::string_view>()(std::string_view(object_begin, size_t(object_end - object_begin)));
    }
    /// Fills the device descriptor of this rule; false when the rule has none.
    virtual bool device_rule(qd_rule &) const { return false; }
    /// Fills the callback table of a rule compiled as a device functor
    /// (quids_amd/user_rule.hpp, device_rule<R>); false when the rule has none.
    virtual bool user_rule(qd_user_rule &) const { return false; }
};

/// An observable evaluated on the device (quids_amd/user_rule.hpp builds one
/// from a functor, device_observable<F>): the sum of obs |m|^2 over a state.
class device_observable_base {
public:
    virtual ~device_observable_base() {}
    virtual double average(qd_ctx *ctx, qd_state *state) const = 0;
};

/// A rule whose four methods run the library's functors (the same code the
/// kernels run) through the C ABI.
class device_rule_base : public rule {
protected:
    qd_rule desc_{};

public:
    bool device_rule(qd_rule &d) const override {
        d = desc_;
        return true;
    }
    void get_num_child(char const *b, char const *e, uint &num_child, uint &max_child_size) const override {
        detail::check(qd_rule_num_child(&desc_, b, uint32_t(e - b), &num_child, &max_child_size), "get_num_child");
    }
    void populate_child(char const *b, char const *e, char *const child, uint const child_id, uint &size,
                        mag_t &mag) const override {
        uint nc, ms;
        get_num_child(b, e, nc, ms);
        double m[2] = {mag.real(), mag.imag()};
        detail::check(qd_rule_populate_child(&desc_, b, uint32_t(e - b), child, ms, child_id, &size, m),
                      "populate_child");
        mag = mag_t(m[0], m[1]);
    }
    size_t hasher(char const *b, char const *e) const override {
        uint64_t h;
        detail::check(qd_rule_hash(&desc_, b, uint32_t(e - b), &h), "hasher");
        return size_t(h);
    }
};

/// A modifier (src/quids.hpp:86): an in-place transform of one object.
/// Built from a device descriptor, or from any host callable (which can be
/// called but not simulated on the GPU).
class modifier_t {
    qd_modifier desc_{};
    bool device_ = false;
    std::function<void(char *, char *, mag_t &)> host_;

public:
    modifier_t() = default;
    explicit modifier_t(const qd_modifier &d) : desc_(d), device_(true) {}
    template <class F, class = decltype(std::declval<F>()((char *)0, (char *)0, std::declval<mag_t &>()))>
    modifier_t(F f) : host_(std::move(f)) {}
    void operator()(char *begin, char *end, mag_t &mag) const {
        if (device_) {
            double m[2] = {mag.real(), mag.imag()};
            detail::check(qd_modifier_apply_host(&desc_, begin, uint32_t(end - begin), m), "modifier");
            mag = mag_t(m[0], m[1]);
        } else if (host_) {
            host_(begin, end, mag);
        }
    }
    bool device_modifier(qd_modifier &d) const {
        d = desc_;
        return device_;
    }
};

/// Wave function (src/quids.hpp:149-335), resident in device memory, with a
/// host copy kept for the read accessors and for appends.
class iteration {
public:
    iteration() { detail::check(qd_state_create(detail::context(), &state_), "qd_state_create"); }
    iteration(char *object_begin_, char *object_end_) : iteration() { append(object_begin_, object_end_); }
    ~iteration() { qd_state_destroy(state_); }
    iteration(const iteration &) = delete;
    iteration &operator=(const iteration &) = delete;

    size_t num_object = 0;
    double total_proba = 1;

    /// iteration::append, src/quids.hpp:174-188
    void append(char const *object_begin_, char const *object_end_, mag_t const mag = 1) {
        pull();
        const size_t size = size_t(object_end_ - object_begin_);
        const size_t off = begin_.back();
        objects_.resize(off + size + get_alignment_offset(uint(size)), 0);
        std::memcpy(objects_.data() + off, object_begin_, size);
        begin_.push_back(objects_.size());
        size_.push_back(uint32_t(size));
        mag_.push_back(mag);
        num_object = size_.size();
        device_valid_ = false;
    }
    /// iteration::pop, src/quids.hpp:194-203
    void pop(size_t n = 1, bool normalize_ = true) {
        if (n < 1) return;
        pull();
        num_object -= n;
        begin_.resize(num_object + 1);
        objects_.resize(begin_.back());
        size_.resize(num_object);
        mag_.resize(num_object);
        device_valid_ = false;
        if (normalize_) normalize();
    }
    /// iteration::normalize, src/quids.hpp:985-1017
    void normalize() {
        push();
        detail::check(qd_normalize(detail::context(), state_, &total_proba), "qd_normalize");
        host_valid_ = false;
    }
    /// iteration::average_value, src/quids.hpp:208-234
    double average_value(const observable_t observable) const {
        pull();
        double avg = 0;
        for (size_t i = 0; i < num_object; ++i)
            avg += observable(objects_.data() + begin_[i], objects_.data() + begin_[i] + size_[i]) * std::norm(mag_[i]);
        return avg;
    }
    /// iteration::average_value for an observable evaluated on the device
    double average_value(const device_observable_base &observable) const {
        push();
        return observable.average(detail::context(), state_);
    }
    /// iteration::get_object (read-only), src/quids.hpp:254-258
    void get_object(size_t const object_id, char const *&object_begin_, uint &object_size_, mag_t &mag) const {
        pull();
        object_size_ = size_[object_id];
        mag = mag_[object_id];
        object_begin_ = objects_.data() + begin_[object_id];
    }
    /// iteration::get_object (read-write), src/quids.hpp:242-246
    void get_object(size_t const object_id, char *&object_begin_, uint &object_size_, mag_t *&mag) {
        pull();
        object_size_ = size_[object_id];
        mag = &mag_[object_id];
        object_begin_ = objects_.data() + begin_[object_id];
        device_valid_ = false;  // the caller may write through the pointers
    }

    qd_state *handle() { return state_; }

private:
    friend void simulate(it_t &, modifier_t const);
    friend void simulate(it_t &, rule_t const *, it_t &, sy_it_t &, size_t, debug_t);

    qd_state *state_ = nullptr;
    mutable std::vector<char> objects_;
    mutable std::vector<uint64_t> begin_{0};
    mutable std::vector<uint32_t> size_;
    mutable std::vector<mag_t> mag_;
    mutable bool host_valid_ = true;
    mutable bool device_valid_ = true;

    void push() const {
        if (device_valid_) return;
        detail::check(qd_state_upload(state_, num_object, objects_.data(), begin_.data(), size_.data(),
                                      reinterpret_cast<const double *>(mag_.data()), total_proba),
                      "qd_state_upload");
        device_valid_ = true;
    }
    void pull() const {
        if (host_valid_) return;
        uint64_t n, bytes;
        double t;
        detail::check(qd_state_info(state_, &n, &bytes, &t), "qd_state_info");
        objects_.resize(bytes);
        begin_.resize(n + 1);
        size_.resize(n);
        mag_.resize(n);
        detail::check(qd_state_download(state_, objects_.data(), begin_.data(), size_.data(),
                                        reinterpret_cast<double *>(mag_.data())),
                      "qd_state_download");
        host_valid_ = true;
    }
    void device_changed() {
        uint64_t n, bytes;
        detail::check(qd_state_info(state_, &n, &bytes, &total_proba), "qd_state_info");
        num_object = n;
        host_valid_ = false;
        device_valid_ = true;
    }
};

/// Symbolic iteration (src/quids.hpp:338-429): the counters of the last
/// step; the step's scratch lives in the device context.
class symbolic_iteration {
public:
    symbolic_iteration() { detail::check(qd_symbolic_create(detail::context(), &sy_), "qd_symbolic_create"); }
    ~symbolic_iteration() { qd_symbolic_destroy(sy_); }
    symbolic_iteration(const symbolic_iteration &) = delete;
    symbolic_iteration &operator=(const symbolic_iteration &) = delete;

    size_t num_object = 0;
    size_t num_object_after_interferences = 0;
    /// counters of the last step (no reference counterpart)
    qd_step_stats stats{};

private:
    friend void simulate(it_t &, rule_t const *, it_t &, sy_it_t &, size_t, debug_t);
    qd_symbolic *sy_ = nullptr;
};

/// quids::simulate(it, modifier), src/quids.hpp:436-438
inline void simulate(it_t &iteration, modifier_t const rule) {
    qd_modifier d;
    if (!rule.device_modifier(d)) throw std::runtime_error("simulate: modifier has no device implementation");
    iteration.push();
    detail::check(qd_apply_modifier(detail::context(), iteration.state_, &d), "qd_apply_modifier");
    iteration.device_changed();
}

/// quids::simulate(it, rule, next, sy_it, max_num_object), src/quids.hpp:448-543.
/// max_num_object == 0: memory-driven truncation; (size_t)-1: no limit.
inline void simulate(it_t &iteration, rule_t const *rule, it_t &next_iteration, sy_it_t &symbolic_iteration,
                     size_t max_num_object = 0, debug_t mid_step_function = [](const char *) {}) {
    qd_rule d;
    qd_user_rule u;
    const bool builtin = rule && rule->device_rule(d);
    if (!builtin && !(rule && rule->user_rule(u)))
        throw std::runtime_error("simulate: rule has no device implementation");
    iteration.push();
    qd_params p = detail::params();
    // the library calls the hook between its phases, with the reference's names
    auto trampoline = [](const char *phase, void *user) { (*static_cast<debug_t *>(user))(phase); };
    detail::check(qd_ctx_set_phase_hook(detail::context(), trampoline, &mid_step_function), "qd_ctx_set_phase_hook");
    const int rc = builtin ? qd_simulate(detail::context(), iteration.state_, &d, next_iteration.state_,
                                         symbolic_iteration.sy_, uint64_t(max_num_object), &p,
                                         &symbolic_iteration.stats)
                           : qd_simulate_user(detail::context(), iteration.state_, &u, next_iteration.state_,
                                              symbolic_iteration.sy_, uint64_t(max_num_object), &p,
                                              &symbolic_iteration.stats);
    qd_ctx_set_phase_hook(detail::context(), nullptr, nullptr);
    detail::check(rc, "qd_simulate");
    symbolic_iteration.num_object = symbolic_iteration.stats.num_candidates;
    symbolic_iteration.num_object_after_interferences = symbolic_iteration.stats.num_unique;
    next_iteration.device_changed();
}

}  // namespace quids

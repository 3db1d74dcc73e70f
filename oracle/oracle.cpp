// oracle.cpp -- CPU restatement of the QuIDS rule-application step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and only to compare against or to time as a CPU baseline.  The
// product path (quids_amd/, include/) never links or calls it.
//
// It restates, in plain C++/OpenMP, what the reference computes:
//   * quids::simulate(it, rule, next, sy_it, max)   src/quids.hpp:448-543
//       compute_num_child                           src/quids.hpp:548-569
//       iteration::truncate (simple truncation)     src/quids.hpp:613-642
//       generate_symbolic_iteration                 src/quids.hpp:647-721
//       compute_collisions                          src/quids.hpp:726-824
//       symbolic_iteration::truncate                src/quids.hpp:866-900
//       finalize                                    src/quids.hpp:905-968
//       normalize                                   src/quids.hpp:985-1017
//   * quids::simulate(it, modifier)                 src/quids.hpp:436-438, 973-980
//   * rules: hadamard and the gates                 src/rules/quantum_computer.hpp:25-75
//            QCGD layout / hash / rules / step      src/rules/qcgd.hpp:11-209, 443-1036
//   * average_value                                 src/quids.hpp:208-234
//
// Where the reference leaves a result unpinned (thread-dependent), this
// restatement fixes one canonical choice, and the HIP path makes the same
// choice, so that the two agree bit for bit:
//   * representative of a hash group = smallest candidate index (the
//     reference keeps the first one inserted, src/quids.hpp:799-806, whose
//     identity depends on the bucket/thread layout);
//   * merged magnitude = pairwise sum of the group's magnitudes taken in
//     ascending candidate index order: adjacent pairs are added level by
//     level ((v0+v1), (v2+v3), ..., an odd last term carried up), which for
//     one to three members is the plain left-to-right sum (the reference sums
//     in insertion order);
//   * top-K truncation orders by |m|^2 descending, then hash ascending (the
//     reference's nth_element breaks ties arbitrarily, src/quids.hpp:884-889);
//     parent pre-truncation breaks ties by ascending object index;
//   * next-state order = ascending representative index (the reference sorts
//     representative indices, src/quids.hpp:927);
//   * total_proba = sum of |m|^2 in a fixed blocked-pairwise order
//     (canonical_sum below; the reference uses an OpenMP reduction);
//   * padding bytes of generated objects are zero (the reference leaves the
//     4 padding bytes of every new sub_node uninitialised, qcgd.hpp:35-39).
//
// Build: g++ -O3 -fopenmp -std=c++20 -ffp-contract=off -shared -fPIC
// (no -march, so no FMA: per-child magnitude products round exactly as in the
// reference build, examples/Makefile:1).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string_view>
#include <unordered_map>
#include <vector>

#include <omp.h>

#include "../include/quids_hip.h"

namespace orc {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

struct cplx {
    double re, im;
};
// (a+bi)(c+di) = (ac-bd) + (ad+bc)i, each product and sum rounded separately,
// which is what GCC emits for std::complex<double> *= (finite operands).
static inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
static inline cplx cneg(cplx a) { return {-a.re, -a.im}; }
static inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
static inline double cnorm(cplx a) { return a.re * a.re + a.im * a.im; }

// ---------------------------------------------------------------------------
// hashing
// ---------------------------------------------------------------------------
// qcgd.hpp:11-25
static inline void hash_combine(u64 &seed, u64 v) {
    const u64 m = 0xc6a4a7935bd1e995ULL;
    seed *= m;
    seed ^= v >> 47;
    seed *= m;
    seed ^= v;
    seed *= m;
    seed += 0xe6546b64ULL;
}

// Default rule hasher, src/quids.hpp:143-145: the library's own
// std::hash<std::string_view> is called directly (no restatement).
static inline u64 std_hash_bytes(const u8 *p, size_t n) {
    return std::hash<std::string_view>()(std::string_view(reinterpret_cast<const char *>(p), n));
}

// ---------------------------------------------------------------------------
// QCGD graph objects, qcgd.hpp:28-112
//   u16 n @0 | bool left[n] @2 | bool right[n] @2+n | u16 name_begin[n+1] @2+2n
//   | sub_node names[m] @4+4n, sub_node = {i16 hmlz_and_element, i16
//   right_or_type, 4 pad bytes, u64 hash} (16 bytes)
// ---------------------------------------------------------------------------
enum : int16_t { DOT_L = -3, DOT_R = -2, ELEMENT = -1 };

struct SubNode {
    int16_t hmlz;
    int16_t type;
    u64 hash;
};

struct Graph {
    std::vector<u8> left, right;
    std::vector<std::vector<SubNode>> names;  // per node, its name tree (root first)
};

static inline u16 rd16(const u8 *p) { u16 v; std::memcpy(&v, p, 2); return v; }
static inline int16_t rdi16(const u8 *p) { int16_t v; std::memcpy(&v, p, 2); return v; }
static inline u64 rd64(const u8 *p) { u64 v; std::memcpy(&v, p, 8); return v; }

static Graph decode(const u8 *obj) {
    Graph g;
    const u16 n = rd16(obj);
    g.left.assign(obj + 2, obj + 2 + n);
    g.right.assign(obj + 2 + n, obj + 2 + 2 * n);
    const u8 *nb = obj + 2 + 2 * n;
    const u8 *subs = obj + 4 + 4 * n;
    g.names.resize(n);
    for (int i = 0; i < n; ++i) {
        const u16 b = rd16(nb + 2 * i), e = rd16(nb + 2 * (i + 1));
        for (u16 k = b; k < e; ++k) {
            const u8 *s = subs + 16 * size_t(k);
            g.names[i].push_back({rdi16(s), rdi16(s + 2), rd64(s + 8)});
        }
    }
    return g;
}

static size_t encoded_size(const Graph &g) {
    size_t m = 0;
    for (auto &v : g.names) m += v.size();
    return 4 + 4 * g.left.size() + 16 * m;
}

static void encode(const Graph &g, u8 *out) {
    const u16 n = u16(g.left.size());
    std::memcpy(out, &n, 2);
    for (int i = 0; i < n; ++i) {
        out[2 + i] = g.left[i] ? 1 : 0;
        out[2 + n + i] = g.right[i] ? 1 : 0;
    }
    u8 *nb = out + 2 + 2 * n;
    u8 *subs = out + 4 + 4 * n;
    u16 off = 0;
    for (int i = 0; i < n; ++i) {
        std::memcpy(nb + 2 * i, &off, 2);
        for (auto &s : g.names[i]) {
            u8 *d = subs + 16 * size_t(off);
            std::memset(d, 0, 16);
            std::memcpy(d, &s.hmlz, 2);
            std::memcpy(d + 2, &s.type, 2);
            std::memcpy(d + 8, &s.hash, 8);
            ++off;
        }
    }
    std::memcpy(nb + 2 * n, &off, 2);
}

// qcgd.hpp:122-146 over the encoded bytes.
static u64 hash_graph(const u8 *obj) {
    const u16 n = rd16(obj);
    const u8 *L = obj + 2, *R = obj + 2 + n, *nb = obj + 2 + 2 * n, *subs = obj + 4 + 4 * n;
    u64 lh = 0, rh = 0, nh = 0;
    for (int i = 0; i < n; ++i) {
        if (L[i]) hash_combine(lh, u64(int64_t(i)));
        if (R[i]) hash_combine(rh, u64(int64_t(i)));
        hash_combine(nh, rd64(subs + 16 * size_t(rd16(nb + 2 * i)) + 8));
    }
    hash_combine(nh, lh);
    hash_combine(nh, rh);
    return nh;
}

// sub_node(node, type) for type in {.l, .r}, qcgd.hpp:43-51.
static SubNode dot_node(const SubNode &root, int16_t type) {
    SubNode s;
    s.type = type;
    s.hmlz = (type == DOT_L && root.hmlz < 0) ? -1 : 1;
    s.hash = root.hash;
    hash_combine(s.hash, u64(int64_t(type)));  // int16 -> size_t sign-extends
    return s;
}

// operations::left / right, qcgd.hpp:194-208.
static std::vector<SubNode> name_left(const std::vector<SubNode> &nm) {
    if (nm[0].type >= 0) return std::vector<SubNode>(nm.begin() + 1, nm.begin() + nm[0].type);
    std::vector<SubNode> r;
    r.push_back(dot_node(nm[0], DOT_L));
    r.insert(r.end(), nm.begin(), nm.end());
    return r;
}
static std::vector<SubNode> name_right(const std::vector<SubNode> &nm) {
    if (nm[0].type >= 0) return std::vector<SubNode>(nm.begin() + nm[0].type, nm.end());
    std::vector<SubNode> r;
    r.push_back(dot_node(nm[0], DOT_R));
    r.insert(r.end(), nm.begin(), nm.end());
    return r;
}
// operations::merge, qcgd.hpp:181-192.
static std::vector<SubNode> name_merge(const std::vector<SubNode> &a, const std::vector<SubNode> &b) {
    if (a[0].type == DOT_L && b[0].type == DOT_R && a.size() > 1 && b.size() > 1 && a[1].hash == b[1].hash)
        return std::vector<SubNode>(a.begin() + 1, a.end());
    SubNode s;
    s.type = int16_t(a.size() + 1);
    s.hmlz = (a[0].hmlz < 0 || b[0].hmlz < 0) ? -1 : 1;
    s.hash = a[0].hash;
    hash_combine(s.hash, b[0].hash);
    std::vector<SubNode> r;
    r.push_back(s);
    r.insert(r.end(), a.begin(), a.end());
    r.insert(r.end(), b.begin(), b.end());
    return r;
}

struct QcgdCoef {
    cplx d, dn, dc, dnc;  // do_, do_not, do_conj, do_not_conj
};

// Forward-mode tangents (the reference gradient the adjoint is checked
// against): a complex value with its derivatives along up to kDir real
// parameters, multiplied through the rules in the same factor order as the
// plain magnitudes.
constexpr int kDir = 8;
struct Dual {
    cplx v;
    cplx t[kDir];
};
static inline Dual cmul(const Dual &a, const Dual &b) {
    Dual r;
    r.v = cmul(a.v, b.v);
    for (int i = 0; i < kDir; ++i) r.t[i] = cadd(cmul(a.t[i], b.v), cmul(a.v, b.t[i]));
    return r;
}
static inline Dual cneg(const Dual &a) {
    Dual r;
    r.v = cneg(a.v);
    for (int i = 0; i < kDir; ++i) r.t[i] = cneg(a.t[i]);
    return r;
}
static inline Dual cscale(const Dual &a, double f) {
    Dual r;
    r.v = cscale(a.v, f);
    for (int i = 0; i < kDir; ++i) r.t[i] = cscale(a.t[i], f);
    return r;
}
struct DualCoef {
    Dual d, dn, dc, dnc;
};

// split_merge::get_num_child, qcgd.hpp:623-642
static u32 split_merge_num_child(const u8 *obj) {
    const u16 n = rd16(obj);
    const u8 *L = obj + 2, *R = obj + 2 + n;
    bool first_split = L[0] && R[0];
    bool last_merge = false;
    if (!first_split && n > 1) last_merge = R[0] && L[n - 1] && !R[n - 1];
    u32 nc = (first_split || last_merge) ? 2 : 1;
    for (int i = int(first_split); i < int(n) - int(last_merge); ++i) {
        bool split = L[i] && R[i];
        bool merge = !split && i + 1 < n && L[i] && R[i + 1] && !L[i + 1];
        if (split || merge) nc *= 2;
    }
    return nc;
}

// split_merge::populate_child, qcgd.hpp:643-848, restated on decoded node
// lists.  Returns the child graph, multiplies *mag in the reference's order.
template <class K, class M>
static Graph split_merge_child(const Graph &g, u32 child_id, const K &k, M *mag) {
    const int n = int(g.left.size());
    const auto &L = g.left, &R = g.right;
    bool first_split = L[0] && R[0];
    bool last_merge = false;
    if (!first_split && n > 1) last_merge = R[0] && L[n - 1] && !R[n - 1];
    u32 cid = child_id;
    first_split = first_split && (cid & 1);
    if (first_split) {
        if (mag) *mag = cmul(*mag, k.d);
        cid >>= 1;
    }
    if (last_merge) {
        if (cid & 1) {
            if (mag) *mag = cmul(*mag, k.dc);
        } else {
            last_merge = false;
            if (mag) *mag = cmul(*mag, cneg(k.dnc));
        }
        cid >>= 1;
    }
    auto op_at = [&](int i, bool &split, bool &merge) {
        split = L[i] && R[i];
        merge = !split && i + 1 < n && L[i] && R[i + 1] && !L[i + 1];
    };
    // magnitudes: node order, first pass of the reference (qcgd.hpp:675-700)
    {
        u32 c = cid;
        for (int i = int(first_split) + int(last_merge); i < n - int(last_merge); ++i) {
            bool split, merge;
            op_at(i, split, merge);
            if (split || merge) {
                if (mag) {
                    if (c & 1) *mag = cmul(*mag, split ? k.d : k.dc);
                    else *mag = cmul(*mag, split ? k.dn : cneg(k.dnc));
                }
                c >>= 1;
            }
        }
    }
    Graph c;
    bool overflow_left = false;  // left(0) goes last
    if (first_split) {
        const auto &nm = g.names[0];
        bool hmlz = true;
        if (nm[0].type >= 0 && nm[1].hmlz > 0) hmlz = false;
        if (hmlz) {
            c.left.push_back(1); c.right.push_back(0); c.names.push_back(name_left(nm));
            c.left.push_back(0); c.right.push_back(1); c.names.push_back(name_right(nm));
        } else {
            overflow_left = true;
            c.left.push_back(0); c.right.push_back(1); c.names.push_back(name_right(nm));
        }
    }
    if (last_merge) {
        c.left.push_back(1); c.right.push_back(1);
        c.names.push_back(name_merge(g.names[n - 1], g.names[0]));
    }
    for (int i = int(first_split) + int(last_merge); i < n - int(last_merge); ++i) {
        bool split, merge;
        op_at(i, split, merge);
        bool done = false;
        if (split || merge) {
            done = cid & 1;
            if (done) {
                if (split) {
                    c.left.push_back(1); c.right.push_back(0); c.names.push_back(name_left(g.names[i]));
                    c.left.push_back(0); c.right.push_back(1); c.names.push_back(name_right(g.names[i]));
                } else {
                    c.left.push_back(1); c.right.push_back(1);
                    c.names.push_back(name_merge(g.names[i], g.names[i + 1]));
                    ++i;
                }
            }
            cid >>= 1;
        }
        if (!done) {
            c.left.push_back(L[i]); c.right.push_back(R[i]); c.names.push_back(g.names[i]);
        }
    }
    if (overflow_left) {
        c.left.push_back(1); c.right.push_back(0); c.names.push_back(name_left(g.names[0]));
    }
    return c;
}

// erase_create (flip where L==R) and coin (flip where L!=R): qcgd.hpp:475-532, 550-605
static u32 flip_num_child(const u8 *obj, bool coin) {
    const u16 n = rd16(obj);
    const u8 *L = obj + 2, *R = obj + 2 + n;
    u32 nc = 1;
    for (int i = 0; i < n; ++i) {
        bool x = bool(L[i]) ^ bool(R[i]);
        if (coin ? x : !x) nc *= 2;
    }
    return nc;
}
template <class K, class M>
static void flip_child(u8 *child, u32 child_id, bool coin, const K &k, M *mag) {
    const u16 n = rd16(child);
    u8 *L = child + 2, *R = child + 2 + n;
    u32 cid = child_id;
    for (int i = 0; i < n; ++i) {
        const bool l = L[i], r = R[i];
        const bool eligible = coin ? (l != r) : (l == r);
        if (!eligible) continue;
        const bool conj = coin ? l : (l && r);
        if (cid & 1) {
            if (mag) *mag = cmul(*mag, conj ? k.dc : k.d);
            L[i] = !l;
            R[i] = !r;
        } else if (mag) {
            *mag = cmul(*mag, conj ? cneg(k.dnc) : k.dn);
        }
        cid >>= 1;
    }
}

// ---------------------------------------------------------------------------
// rule dispatch
// ---------------------------------------------------------------------------
struct Rule {
    qd_rule d;
    QcgdCoef k() const {
        return {{d.do_re, d.do_im}, {d.do_not_re, d.do_not_im}, {d.do_conj_re, d.do_conj_im},
                {d.do_not_conj_re, d.do_not_conj_im}};
    }
};

static bool num_child(const qd_rule &r, const u8 *p, u32 psize, u32 &nc, u32 &max_size) {
    switch (r.kind) {
    case QD_RULE_HADAMARD:
        nc = 2; max_size = psize; return true;
    case QD_RULE_QCGD_ERASE_CREATE:
        nc = flip_num_child(p, false); max_size = psize; return true;
    case QD_RULE_QCGD_COIN:
        nc = flip_num_child(p, true); max_size = psize; return true;
    case QD_RULE_QCGD_SPLIT_MERGE:
        nc = split_merge_num_child(p); max_size = 4 * psize; return true;
    }
    return false;
}

// populate_child: writes the child into out (capacity >= max child size),
// returns its size; multiplies *mag (may be null).
template <class K, class M>
static u32 populate_t(const qd_rule &r, const K &k, const u8 *p, u32 psize, u32 child_id, u8 *out, M *mag) {
    switch (r.kind) {
    case QD_RULE_HADAMARD: {
        const u8 b = p[r.bit];
        if (mag) *mag = cscale(*mag, (b && child_id) ? -r.coef : r.coef);
        std::memcpy(out, p, psize);
        out[r.bit] ^= u8(!child_id);
        return psize;
    }
    case QD_RULE_QCGD_ERASE_CREATE:
    case QD_RULE_QCGD_COIN:
        std::memcpy(out, p, psize);
        flip_child(out, child_id, r.kind == QD_RULE_QCGD_COIN, k, mag);
        return psize;
    case QD_RULE_QCGD_SPLIT_MERGE: {
        Graph g = decode(p);
        Graph c = split_merge_child(g, child_id, k, mag);
        encode(c, out);
        return u32(encoded_size(c));
    }
    }
    return 0;
}
static u32 populate(const qd_rule &r, const u8 *p, u32 psize, u32 child_id, u8 *out, cplx *mag) {
    return populate_t(r, Rule{r}.k(), p, psize, child_id, out, mag);
}
// dcoef[t]: rule whose coefficient fields are the derivatives of r's
// coefficients along parameter t (t < D).
static DualCoef dual_coef(const qd_rule &r, const qd_rule *dcoef, int D) {
    DualCoef k{};
    const QcgdCoef v = Rule{r}.k();
    k.d.v = v.d;
    k.dn.v = v.dn;
    k.dc.v = v.dc;
    k.dnc.v = v.dnc;
    for (int t = 0; t < D; ++t) {
        const QcgdCoef d = Rule{dcoef[t]}.k();
        k.d.t[t] = d.d;
        k.dn.t[t] = d.dn;
        k.dc.t[t] = d.dc;
        k.dnc.t[t] = d.dnc;
    }
    return k;
}

static u64 rule_hash(const qd_rule &r, const u8 *obj, u32 size) {
    if (r.kind == QD_RULE_HADAMARD) return std_hash_bytes(obj, size);
    return hash_graph(obj);
}

static bool apply_modifier(const qd_modifier &m, u8 *obj, u32 size, cplx &mag) {
    switch (m.kind) {
    case QD_MOD_CNOT:  // quantum_computer.hpp:25-29
        obj[m.bit] ^= obj[m.control_bit];
        return true;
    case QD_MOD_XGATE:  // 52-56
        obj[m.bit] = !obj[m.bit];
        return true;
    case QD_MOD_YGATE:  // 58-66: mag *= 1.0i; if bit: mag *= -1; flip
        mag = cmul(mag, cplx{0.0, 1.0});
        if (obj[m.bit]) mag = cscale(mag, -1.0);
        obj[m.bit] = !obj[m.bit];
        return true;
    case QD_MOD_ZGATE:  // 68-75 (flips the bit too)
        if (obj[m.bit]) mag = cscale(mag, -1.0);
        obj[m.bit] = !obj[m.bit];
        return true;
    case QD_MOD_QCGD_STEP:  // qcgd.hpp:443-449
    case QD_MOD_QCGD_REVERSED_STEP: {  // 451-457
        const u16 n = rd16(obj);
        u8 *L = obj + 2, *R = obj + 2 + n;
        if (n == 0) return true;
        if (m.kind == QD_MOD_QCGD_STEP) {
            std::rotate(L, L + 1, L + n);
            std::rotate(R, R + n - 1, R + n);
        } else {
            std::rotate(R, R + 1, R + n);
            std::rotate(L, L + n - 1, L + n);
        }
        return true;
    }
    }
    return false;
}

// ---------------------------------------------------------------------------
// canonical sum of non-negative doubles (the order the HIP normalize kernel
// uses): chunks of 2048 in index order; inside a chunk, 256 lanes each sum
// elements t, t+256, ..., t+1792 in that order, then a halving tree
// s[t] += s[t+stride] for stride = 128, 64, ..., 1; chunk sums are reduced
// the same way until one value is left.
// ---------------------------------------------------------------------------
static double canonical_sum(std::vector<double> v) {
    if (v.empty()) return 0.0;
    while (v.size() > 1) {
        const size_t nchunk = (v.size() + 2047) / 2048;
        std::vector<double> nx(nchunk);
        for (size_t c = 0; c < nchunk; ++c) {
            double s[256];
            for (int t = 0; t < 256; ++t) {
                double a = 0.0;
                for (int j = 0; j < 8; ++j) {
                    size_t idx = c * 2048 + size_t(t) + 256 * size_t(j);
                    if (idx < v.size()) a += v[idx];
                }
                s[t] = a;
            }
            for (int stride = 128; stride >= 1; stride /= 2)
                for (int t = 0; t < stride; ++t) s[t] = s[t] + s[t + stride];
            nx[c] = s[0];
        }
        v.swap(nx);
    }
    return v[0];
}

// Pairwise sum in the given order: level by level, adjacent pairs are
// added and an odd last term is carried to the next level.
static cplx pairwise_sum(std::vector<cplx> v) {
    while (v.size() > 1) {
        std::vector<cplx> w((v.size() + 1) / 2);
        for (size_t j = 0; j < w.size(); ++j) w[j] = 2 * j + 1 < v.size() ? cadd(v[2 * j], v[2 * j + 1]) : v[2 * j];
        v.swap(w);
    }
    return v.empty() ? cplx{0, 0} : v[0];
}

struct State {
    std::vector<u8> objects;
    std::vector<u64> begin{0};
    std::vector<u32> size;
    std::vector<cplx> mag;
    double total_proba = 1;
    size_t n() const { return size.size(); }
};

// Probabilistic truncation (src/quids.hpp:594-608, 829-845, 884-895): keep
// the smallest u / |m|^2, i.e. the largest |m|^2 / u.  The reference draws u
// from per-thread generators seeded by rand() (random.hpp:12-15); this
// restatement, like the library, draws it from Philox2x64-10 (Salmon et
// al., SC 2011: 10 rounds of (c0, c1) <- (hi(M c0) ^ k ^ c1, lo(M c0)),
// k += W) with counter (hash, 0) and key = seed, and takes the top 53 bits
// as u in (0, 1].
static void philox2x64_10(u64 &c0, u64 &c1, u64 k) {
    for (int r = 0; r < 10; ++r) {
        const unsigned __int128 p = (unsigned __int128)0xD2B74407B1CE6E93ull * c0;
        const u64 hi = u64(p >> 64), lo = u64(p);
        c0 = hi ^ k ^ c1;
        c1 = lo;
        k += 0x9E3779B97F4A7C15ull;
    }
}
static double trunc_uniform(u64 hash, u64 seed) {
    u64 c0 = hash, c1 = 0;
    philox2x64_10(c0, c1, seed);
    return double((c0 >> 11) + 1) * 0x1p-53;
}
// the value whose larger entries are kept
static double trunc_key(cplx m, u64 hash, const qd_params &prm) {
    const double v = cnorm(m);
    return prm.simple_truncation ? v : v / trunc_uniform(hash, prm.seed);
}

static inline u32 align_pad(u32 size, u32 align) {
    if (align <= 1) return 0;
    u32 off = align - size % align;
    return off == align ? 0 : off;
}

// ---------------------------------------------------------------------------
// simulate, src/quids.hpp:448-543 (simple truncation, explicit max)
// ---------------------------------------------------------------------------
// With D > 0, in_tan holds D tangents per input object (object-major) and
// out_tan receives the output's: the forward-mode derivative of the step
// along D real parameters, dcoef[t] holding the derivatives of the rule's
// coefficients along parameter t.
// max_survivors caps the uniques as max_num_object caps the parents
// (has_survivor_cap false: the same value); the reference uses one value for
// both unless its memory-driven searches chose them (src/quids.hpp:472-485,
// 523-535).
static int simulate(const State &in, const qd_rule &rule, State &out, u64 max_num_object,
                    const qd_params &prm, qd_step_stats &st, const qd_rule *dcoef = nullptr, int D = 0,
                    const std::vector<cplx> *in_tan = nullptr, std::vector<cplx> *out_tan = nullptr,
                    bool has_survivor_cap = false, u64 max_survivors = 0, bool normalize = true) {
    if (!has_survivor_cap) max_survivors = max_num_object;
    const size_t N = in.n();
    std::memset(&st, 0, sizeof st);
    st.num_parent = N;

    std::vector<u32> nchild(N);
    bool ok = true;
#pragma omp parallel for reduction(&& : ok)
    for (size_t p = 0; p < N; ++p) {
        u32 ms;
        ok = num_child(rule, in.objects.data() + in.begin[p], in.size[p], nchild[p], ms) && ok;
    }
    if (!ok) return QD_ERR_UNSUPPORTED;

    // parent pre-truncation (src/quids.hpp:613-642): top max by |m|^2, ties by index
    std::vector<u64> kept;
    if (max_num_object < N) {
        std::vector<u64> ord(N);
        for (size_t p = 0; p < N; ++p) ord[p] = p;
        std::vector<double> key(N);
        for (size_t p = 0; p < N; ++p)
            key[p] = trunc_key(in.mag[p],
                               prm.simple_truncation ? 0 : rule_hash(rule, in.objects.data() + in.begin[p], in.size[p]),
                               prm);
        std::nth_element(ord.begin(), ord.begin() + max_num_object, ord.end(), [&](u64 a, u64 b) {
            return key[a] != key[b] ? key[a] > key[b] : a < b;
        });
        ord.resize(max_num_object);
        std::sort(ord.begin(), ord.end());
        kept = ord;
    } else {
        kept.resize(N);
        for (size_t p = 0; p < N; ++p) kept[p] = p;
    }
    st.num_parent_kept = kept.size();

    // candidate index space: kept parents ascending, then child_id
    std::vector<u64> cbegin(kept.size() + 1, 0);
    for (size_t i = 0; i < kept.size(); ++i) cbegin[i + 1] = cbegin[i] + nchild[kept[i]];
    const u64 C = cbegin.back();
    st.num_candidates = C;

    std::vector<u64> H(C);
    std::vector<cplx> M(C);
    std::vector<cplx> T(D ? C * D : 0);
    std::vector<u32> pidx(C), cid(C);
    const DualCoef dk = D ? dual_coef(rule, dcoef, D) : DualCoef{};
#pragma omp parallel
    {
        std::vector<u8> buf;
#pragma omp for schedule(dynamic, 64)
        for (size_t i = 0; i < kept.size(); ++i) {
            const u64 p = kept[i];
            const u8 *po = in.objects.data() + in.begin[p];
            u32 nc, ms;
            num_child(rule, po, in.size[p], nc, ms);
            if (buf.size() < ms + 64) buf.resize(ms + 64);
            for (u32 j = 0; j < nchild[p]; ++j) {
                const u64 c = cbegin[i] + j;
                u32 sz;
                if (D) {
                    Dual m{};
                    m.v = in.mag[p];
                    for (int t = 0; t < D; ++t) m.t[t] = (*in_tan)[p * D + t];
                    sz = populate_t(rule, dk, po, in.size[p], j, buf.data(), &m);
                    M[c] = m.v;
                    for (int t = 0; t < D; ++t) T[c * D + t] = m.t[t];
                } else {
                    cplx m = in.mag[p];
                    sz = populate(rule, po, in.size[p], j, buf.data(), &m);
                    M[c] = m;
                }
                H[c] = rule_hash(rule, buf.data(), sz);
                pidx[c] = u32(i);
                cid[c] = j;
            }
        }
    }

    // interference (src/quids.hpp:726-824): group by hash, bucketed by the
    // top hash bits; inside a bucket candidates are visited in ascending index
    // so the first one seen is the representative and sums run in order.
    const int bucket_bits = 10;
    const size_t B = size_t(1) << bucket_bits;
    std::vector<u64> bcount(B + 1, 0);
    for (u64 c = 0; c < C; ++c) ++bcount[(H[c] >> (64 - bucket_bits)) + 1];
    for (size_t b = 0; b < B; ++b) bcount[b + 1] += bcount[b];
    std::vector<u64> order(C);
    {
        std::vector<u64> cur(bcount.begin(), bcount.end() - 1);
        for (u64 c = 0; c < C; ++c) order[cur[H[c] >> (64 - bucket_bits)]++] = c;
    }
    struct Uniq {
        u64 hash, rep;
        cplx mag;
        std::vector<cplx> tan;
    };
    std::vector<std::vector<Uniq>> per_bucket(B);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t b = 0; b < B; ++b) {
        std::unordered_map<u64, size_t> idx;
        std::vector<std::vector<cplx>> members;
        std::vector<std::vector<u64>> mc;
        auto &v = per_bucket[b];
        idx.reserve(bcount[b + 1] - bcount[b]);
        for (u64 k = bcount[b]; k < bcount[b + 1]; ++k) {
            const u64 c = order[k];
            auto [it, fresh] = idx.try_emplace(H[c], v.size());
            if (fresh) {
                v.push_back({H[c], c, M[c], {}});
                members.emplace_back();
                mc.emplace_back();
            }
            members[it->second].push_back(M[c]);  // ascending candidate order
            mc[it->second].push_back(c);
        }
        for (size_t g = 0; g < v.size(); ++g) {
            v[g].mag = pairwise_sum(members[g]);
            for (int t = 0; t < D; ++t) {
                std::vector<cplx> tv;
                for (u64 c : mc[g]) tv.push_back(T[c * D + t]);
                v[g].tan.push_back(pairwise_sum(tv));
            }
        }
    }
    std::vector<Uniq> uniq;
    for (auto &v : per_bucket)
        for (auto &u : v)
            if (cnorm(u.mag) > prm.tolerance) uniq.push_back(u);
    st.num_unique = uniq.size();

    // truncation (src/quids.hpp:866-900): top max by |m|^2 desc, hash asc
    if (max_survivors < uniq.size()) {
        std::nth_element(uniq.begin(), uniq.begin() + max_survivors, uniq.end(),
                         [&](const Uniq &a, const Uniq &b) {
                             const double na = trunc_key(a.mag, a.hash, prm), nb = trunc_key(b.mag, b.hash, prm);
                             return na != nb ? na > nb : a.hash < b.hash;
                         });
        uniq.resize(max_survivors);
    }
    // finalize (src/quids.hpp:905-968): ascending representative order
    std::sort(uniq.begin(), uniq.end(), [](const Uniq &a, const Uniq &b) { return a.rep < b.rep; });
    const size_t K = uniq.size();
    st.num_survivors = K;

    out = State();
    out.size.resize(K);
    out.mag.resize(K);
    out.begin.assign(K + 1, 0);
    std::vector<std::vector<u8>> objs(K);
#pragma omp parallel
    {
        std::vector<u8> buf;
#pragma omp for schedule(dynamic, 64)
        for (size_t k = 0; k < K; ++k) {
            const u64 c = uniq[k].rep;
            const u64 p = kept[pidx[c]];
            const u8 *po = in.objects.data() + in.begin[p];
            u32 nc, ms;
            num_child(rule, po, in.size[p], nc, ms);
            if (buf.size() < ms + 64) buf.resize(ms + 64);
            u32 sz = populate(rule, po, in.size[p], cid[c], buf.data(), nullptr);
            objs[k].assign(buf.begin(), buf.begin() + sz);
            out.size[k] = sz;
            out.mag[k] = uniq[k].mag;
        }
    }
    for (size_t k = 0; k < K; ++k) out.begin[k + 1] = out.begin[k] + out.size[k] + align_pad(out.size[k], prm.align_byte_length);
    out.objects.assign(out.begin[K], 0);
    for (size_t k = 0; k < K; ++k) std::memcpy(out.objects.data() + out.begin[k], objs[k].data(), objs[k].size());

    // normalize (src/quids.hpp:985-1017)
    std::vector<double> nv(K);
    for (size_t k = 0; k < K; ++k) nv[k] = cnorm(out.mag[k]);
    out.total_proba = canonical_sum(nv);
    const double f = std::sqrt(out.total_proba);
    if (D && out_tan) {
        // d(y / |y|) = (dy - y Re<y, dy> / |y|^2) / |y|
        out_tan->assign(K * D, cplx{0, 0});
        for (int t = 0; t < D; ++t) {
            double sig = 0;
            for (size_t k = 0; k < K; ++k)
                sig += out.mag[k].re * uniq[k].tan[t].re + out.mag[k].im * uniq[k].tan[t].im;
            for (size_t k = 0; k < K; ++k) {
                const cplx d = uniq[k].tan[t];
                (*out_tan)[k * D + t] = {(d.re - out.mag[k].re * sig / out.total_proba) / f,
                                         (d.im - out.mag[k].im * sig / out.total_proba) / f};
            }
        }
    }
    if (K > 0 && f != 1 && normalize)
        for (auto &m : out.mag) m = {m.re / f, m.im / f};
    if (K == 0) out.total_proba = 0;
    st.total_proba = out.total_proba;
    st.output_bytes = out.begin[K];
    return QD_OK;
}

}  // namespace orc

using namespace orc;

extern "C" {

struct or_state {
    State s;
    int D = 0;               // tangent directions carried (forward mode)
    std::vector<cplx> tan;   // D per object, object-major
};

or_state *or_state_new(uint64_t n, const void *objects, const uint64_t *begin, const uint32_t *size,
                       const double *mag, double total_proba) {
    auto *st = new or_state;
    const u8 *o = static_cast<const u8 *>(objects);
    st->s.objects.assign(o, o + (n ? begin[n] : 0));
    st->s.begin.assign(begin, begin + n + 1);
    st->s.size.assign(size, size + n);
    st->s.mag.resize(n);
    for (uint64_t i = 0; i < n; ++i) st->s.mag[i] = {mag[2 * i], mag[2 * i + 1]};
    st->s.total_proba = total_proba;
    return st;
}

void or_state_free(or_state *s) { delete s; }

// OpenMP threads of the calling thread's later calls (per host thread, so
// that several host threads can each run small checks single-threaded).
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void or_state_info(const or_state *s, uint64_t *n, uint64_t *bytes, double *total) {
    *n = s->s.n();
    *bytes = s->s.begin.back();
    *total = s->s.total_proba;
}

void or_state_get(const or_state *s, void *objects, uint64_t *begin, uint32_t *size, double *mag) {
    const size_t n = s->s.n();
    if (objects) std::memcpy(objects, s->s.objects.data(), s->s.begin.back());
    if (begin) std::memcpy(begin, s->s.begin.data(), 8 * (n + 1));
    if (size) std::memcpy(size, s->s.size.data(), 4 * n);
    if (mag)
        for (size_t i = 0; i < n; ++i) {
            mag[2 * i] = s->s.mag[i].re;
            mag[2 * i + 1] = s->s.mag[i].im;
        }
}

int or_simulate(const or_state *in, const qd_rule *rule, or_state **out, uint64_t max_num_object,
                const qd_params *prm, qd_step_stats *st) {
    auto *o = new or_state;
    int rc = simulate(in->s, *rule, o->s, max_num_object, *prm, *st);
    if (rc) {
        delete o;
        *out = nullptr;
        return rc;
    }
    *out = o;
    return QD_OK;
}

// The step with separate truncation counts for the parents and the uniques
// (the counts of the reference's memory-driven searches,
// src/quids.hpp:472-485, 523-535); with normalize 0 the magnitudes are left
// as the interference summed them (the values the truncation orders).
int or_simulate_caps(const or_state *in, const qd_rule *rule, or_state **out, uint64_t max_parents,
                     uint64_t max_survivors, const qd_params *prm, qd_step_stats *st, int normalize) {
    auto *o = new or_state;
    int rc = simulate(in->s, *rule, o->s, max_parents, *prm, *st, nullptr, 0, nullptr, nullptr, true, max_survivors,
                      normalize != 0);
    if (rc) {
        delete o;
        *out = nullptr;
        return rc;
    }
    *out = o;
    return QD_OK;
}

int or_modifier(or_state *s, const qd_modifier *m) {
    bool ok = true;
    const size_t n = s->s.n();
#pragma omp parallel for reduction(&& : ok)
    for (size_t i = 0; i < n; ++i) {
        u8 *o = s->s.objects.data() + s->s.begin[i];
        if (s->D) {  // tangents take the factor the modifier applies to the magnitude
            std::vector<u8> tmp(o, o + s->s.size[i]);
            cplx f{1, 0};
            apply_modifier(*m, tmp.data(), s->s.size[i], f);
            for (int t = 0; t < s->D; ++t) s->tan[i * s->D + t] = cmul(s->tan[i * s->D + t], f);
        }
        ok = apply_modifier(*m, o, s->s.size[i], s->s.mag[i]) && ok;
    }
    return ok ? QD_OK : QD_ERR_UNSUPPORTED;
}

// Forward-mode tangents: set / get D tangents per object (re, im pairs,
// object-major), and a step carrying them (dcoef: D derivative rules).
int or_state_set_tangent(or_state *s, int D, const double *tan) {
    if (D < 0 || D > kDir) return QD_ERR_INVALID;
    s->D = D;
    s->tan.resize(s->s.n() * D);
    for (size_t i = 0; i < s->tan.size(); ++i) s->tan[i] = {tan[2 * i], tan[2 * i + 1]};
    return QD_OK;
}

void or_state_get_tangent(const or_state *s, double *tan) {
    for (size_t i = 0; i < s->tan.size(); ++i) {
        tan[2 * i] = s->tan[i].re;
        tan[2 * i + 1] = s->tan[i].im;
    }
}

int or_simulate_jvp(const or_state *in, const qd_rule *rule, const qd_rule *dcoef, or_state **out,
                    uint64_t max_num_object, const qd_params *prm, qd_step_stats *st) {
    auto *o = new or_state;
    o->D = in->D;
    int rc = simulate(in->s, *rule, o->s, max_num_object, *prm, *st, dcoef, in->D, &in->tan, &o->tan);
    if (rc) {
        delete o;
        *out = nullptr;
        return rc;
    }
    *out = o;
    return QD_OK;
}

int or_num_child(const qd_rule *r, const void *obj, uint32_t size, uint32_t *nc, uint32_t *max_size) {
    return num_child(*r, static_cast<const u8 *>(obj), size, *nc, *max_size) ? QD_OK : QD_ERR_UNSUPPORTED;
}

int or_populate(const qd_rule *r, const void *obj, uint32_t size, uint32_t child_id, void *out,
                uint32_t *out_size, double *mag) {
    cplx m{mag ? mag[0] : 1.0, mag ? mag[1] : 0.0};
    *out_size = populate(*r, static_cast<const u8 *>(obj), size, child_id, static_cast<u8 *>(out), &m);
    if (mag) {
        mag[0] = m.re;
        mag[1] = m.im;
    }
    return QD_OK;
}

uint64_t or_hash(const qd_rule *r, const void *obj, uint32_t size) {
    return rule_hash(*r, static_cast<const u8 *>(obj), size);
}

uint64_t or_std_hash(const void *bytes, uint64_t n) { return std_hash_bytes(static_cast<const u8 *>(bytes), n); }

uint64_t or_hash_combine(uint64_t seed, uint64_t v) {
    hash_combine(seed, v);
    return seed;
}

void or_philox2x64_10(uint64_t c0, uint64_t c1, uint64_t key, uint64_t *out) {
    philox2x64_10(c0, c1, key);
    out[0] = c0;
    out[1] = c1;
}

double or_canonical_sum(const double *v, uint64_t n) { return canonical_sum(std::vector<double>(v, v + n)); }

// iteration::average_value (src/quids.hpp:208-234) for the QCGD observables
// the adjoint loss uses: 0 = num_nodes, 1 = particle count sum(L+R),
// 2 = [num_nodes == arg].  Summed sequentially in object order.
static double observable_of(const u8 *o, int observable, int arg) {
    const u16 n = rd16(o);
    if (observable == 0) return n;
    if (observable == 2) return (n == arg) ? 1.0 : 0.0;
    double v = 0;
    for (int k = 0; k < n; ++k) v += double(o[2 + k] != 0) + double(o[2 + n + k] != 0);
    return v;
}

// d average_value / d parameter t = sum obs 2 Re(conj(m) dm_t), into dv[t].
void or_average_value_jvp(const or_state *s, int observable, int arg, double *dv) {
    for (int t = 0; t < s->D; ++t) dv[t] = 0;
    for (size_t i = 0; i < s->s.n(); ++i) {
        const double v = observable_of(s->s.objects.data() + s->s.begin[i], observable, arg);
        const cplx m = s->s.mag[i];
        for (int t = 0; t < s->D; ++t) {
            const cplx d = s->tan[i * s->D + t];
            dv[t] += 2 * v * (m.re * d.re + m.im * d.im);
        }
    }
}

double or_average_value(const or_state *s, int observable, int arg) {
    double avg = 0;
    for (size_t i = 0; i < s->s.n(); ++i) {
        const u8 *o = s->s.objects.data() + s->s.begin[i];
        const u16 n = rd16(o);
        double v = 0;
        if (observable == 0) v = n;
        else if (observable == 1) {
            for (int k = 0; k < n; ++k) v += double(o[2 + k] != 0) + double(o[2 + n + k] != 0);
        } else v = (n == arg) ? 1.0 : 0.0;
        avg += v * cnorm(s->s.mag[i]);
    }
    return avg;
}

}  // extern "C"

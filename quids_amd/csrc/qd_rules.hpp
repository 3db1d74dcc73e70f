// qd_rules.hpp -- QuIDS rules as __host__ __device__ functors.
//
// One implementation per rule serves the device kernels (hot path) and the
// host-side single-object calls of the C ABI (qd_rule_num_child / populate /
// hash), which back the virtual methods of the C++ header's rule classes.
//
// Reference entities restated here:
//   quids::rule (num_child / populate_child / populate_child_simple / hasher)
//                                           src/quids.hpp:105-146
//   hadamard, cnot, Xgate, Ygate, Zgate     src/rules/quantum_computer.hpp:25-75
//   QCGD graph layout and hash_combine      src/rules/qcgd.hpp:11-112, 122-146
//   erase_create, coin, split_merge         src/rules/qcgd.hpp:459-1036
//   step / reversed_step                    src/rules/qcgd.hpp:443-457
//   default hasher std::hash<string_view>   src/quids.hpp:143-145
//
// Hot-path specialisation: the reference materialises every child and hashes
// its bytes (src/quids.hpp:711-718).  QCGD child hashes depend only on the
// root hash of each child node's name tree and on the child's particle bits
// (qcgd.hpp:122-146), so here a per-parent table holds, for every node, the
// root hash of the node itself, of its split halves and of its merge with
// the next node (qcgd.hpp:181-208), and a child hash is a fold over the
// child's node sequence read from that table.  Bytes are only produced for
// the survivors (finalize).
//
// All floating-point products keep the reference's operand order and are
// compiled with -ffp-contract=off, so every child magnitude is bit-identical
// to the reference's (which is built without FMA, examples/Makefile:1).
#pragma once

#include <stdint.h>
#include <string.h>

#include "../../include/quids_hip.h"

#if defined(__HIPCC__)
#define QD_HD __host__ __device__ inline
#else
#define QD_HD inline
#endif

namespace qd {

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int16_t i16;

struct cplx {
    double re, im;
};

QD_HD cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
QD_HD cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
QD_HD cplx cneg(cplx a) { return {-a.re, -a.im}; }
QD_HD double cnorm(cplx a) { return a.re * a.re + a.im * a.im; }

// ---------------------------------------------------------------------------
// unaligned little-endian access (objects are byte strings; QCGD sub_node
// hashes sit at 4-byte-aligned addresses when the node count is even)
// ---------------------------------------------------------------------------
QD_HD u16 ld16(const u8 *p) { return u16(p[0]) | u16(u16(p[1]) << 8); }
QD_HD i16 ldi16(const u8 *p) { return i16(ld16(p)); }
QD_HD u64 ld64(const u8 *p) {
    u64 v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
QD_HD void st16(u8 *p, u16 v) {
    p[0] = u8(v);
    p[1] = u8(v >> 8);
}

// ---------------------------------------------------------------------------
// hashes
// ---------------------------------------------------------------------------
constexpr u64 kMul = 0xc6a4a7935bd1e995ULL;

// utils::hash_combine, qcgd.hpp:11-25
QD_HD u64 hash_combine(u64 seed, u64 v) {
    seed *= kMul;
    seed ^= v >> 47;
    seed *= kMul;
    seed ^= v;
    seed *= kMul;
    seed += 0xe6546b64ULL;
    return seed;
}
// Same for v < 2^47 (node indices): v >> 47 is zero, so the first two
// multiplications fold into one by kMul^2 mod 2^64.
constexpr u64 kMul2 = kMul * kMul;
QD_HD u64 hash_combine_small(u64 seed, u64 v) {
    seed *= kMul2;
    seed ^= v;
    seed *= kMul;
    seed += 0xe6546b64ULL;
    return seed;
}

// Folds carried in pre-multiplied form (the fast walks' child hashes).
// hash_combine starts by multiplying its seed by kMul, hash_combine_small by
// kMul^2; a fold can instead carry u = seed * kMul (or seed * kMul^2) and
// end each step with one multiplication by a combined constant, which drops
// one of hash_combine's three 64-bit multiplications and one of
// hash_combine_small's two.  The seed comes back as u times the inverse of
// the constant mod 2^64 (kMul is odd), once at the end.  The values are
// those of the plain folds, bit for bit: only the association of the
// multiplications mod 2^64 changes.
constexpr u64 kAdd = 0xe6546b64ULL;
constexpr u64 kMul3 = kMul2 * kMul;
constexpr u64 inverse_mod_2_64(u64 a) {  // a odd: Newton's iteration doubles the correct low bits
    u64 x = a;                          // a * a == 1 mod 8
    for (int i = 0; i < 5; ++i) x *= 2 - a * x;
    return x;
}
constexpr u64 kInvMul = inverse_mod_2_64(kMul);
constexpr u64 kInvMul2 = inverse_mod_2_64(kMul2);
static_assert(kMul * kInvMul == 1 && kMul2 * kInvMul2 == 1, "inverses mod 2^64");
// u = s * kMul^2  ->  hash_combine_small(s, v) * kMul^2
QD_HD u64 small_step_pm(u64 u, u64 v) { return (u ^ v) * kMul3 + kAdd * kMul2; }
// u = s * kMul  ->  hash_combine(s, v) * kMul
QD_HD u64 step_pm(u64 u, u64 v) { return (((u ^ (v >> 47)) * kMul) ^ v) * (kMul * kMul) + kAdd * kMul; }
// u = s * kMul  ->  hash_combine(s, v)
QD_HD u64 last_step_pm(u64 u, u64 v) { return (((u ^ (v >> 47)) * kMul) ^ v) * kMul + kAdd; }
// The child hash hash_combine(hash_combine(s, lh), rh) (qcgd.hpp:143-145)
// from the pre-multiplied s * kMul, lh * kMul^2 and rh * kMul^2.
QD_HD u64 finish_pm(u64 su, u64 lu, u64 ru) { return last_step_pm(step_pm(su, lu * kInvMul2), ru * kInvMul2); }

// Default rule hasher: std::hash<std::string_view> of libstdc++ on LP64,
// which is MurmurHash64A (Austin Appleby's published algorithm) with seed
// 0xc70f6907 and the trailing bytes loaded little-endian.  Checked against
// the host library's std::hash in the tests.
QD_HD u64 murmur64a_shift_mix(u64 v) { return v ^ (v >> 47); }
QD_HD u64 hash_bytes(const u8 *p, u32 len) {
    const u64 seed = 0xc70f6907ULL;
    u64 h = seed ^ (u64(len) * kMul);
    const u32 nblk = len / 8;
    for (u32 b = 0; b < nblk; ++b) {
        u64 k = 0;
        for (int i = 7; i >= 0; --i) k = (k << 8) | p[8 * b + i];
        k = murmur64a_shift_mix(k * kMul) * kMul;
        h ^= k;
        h *= kMul;
    }
    const u32 rem = len & 7;
    if (rem) {
        u64 t = 0;
        for (int i = int(rem) - 1; i >= 0; --i) t = (t << 8) | p[8 * nblk + i];
        h ^= t;
        h *= kMul;
    }
    h = murmur64a_shift_mix(h) * kMul;
    return murmur64a_shift_mix(h);
}
// hash_bytes of the object p with byte `pos` xor-ed with `flip`, without
// materialising it (hadamard children).
QD_HD u64 hash_bytes_flip(const u8 *p, u32 len, u32 pos, u8 flip) {
    const u64 seed = 0xc70f6907ULL;
    u64 h = seed ^ (u64(len) * kMul);
    const u32 nblk = len / 8;
    for (u32 b = 0; b < nblk; ++b) {
        u64 k = 0;
        for (int i = 7; i >= 0; --i) {
            u8 c = p[8 * b + i];
            if (8 * b + u32(i) == pos) c ^= flip;
            k = (k << 8) | c;
        }
        k = murmur64a_shift_mix(k * kMul) * kMul;
        h ^= k;
        h *= kMul;
    }
    const u32 rem = len & 7;
    if (rem) {
        u64 t = 0;
        for (int i = int(rem) - 1; i >= 0; --i) {
            u8 c = p[8 * nblk + i];
            if (8 * nblk + u32(i) == pos) c ^= flip;
            t = (t << 8) | c;
        }
        h ^= t;
        h *= kMul;
    }
    h = murmur64a_shift_mix(h) * kMul;
    return murmur64a_shift_mix(h);
}

// ---------------------------------------------------------------------------
// QCGD graph layout, qcgd.hpp:28-112
//   u16 n @0 | bool left[n] @2 | bool right[n] @2+n | u16 name_begin[n+1]
//   @2+2n | sub_node names[m] @4+4n; sub_node = {i16 hmlz_and_element,
//   i16 right_or_type, 4 padding bytes, u64 hash} (16 bytes).
// right_or_type: -3 ".l", -2 ".r", -1 element, >= 0 offset of the right half
// of a pair node.
// ---------------------------------------------------------------------------
enum : i16 { QCGD_DOT_L = -3, QCGD_DOT_R = -2, QCGD_ELEMENT = -1 };

struct Graph {
    const u8 *o;
    u32 n;
    QD_HD explicit Graph(const u8 *obj) : o(obj), n(ld16(obj)) {}
    QD_HD bool L(u32 i) const { return o[2 + i] != 0; }
    QD_HD bool R(u32 i) const { return o[2 + n + i] != 0; }
    QD_HD u32 name_begin(u32 i) const { return ld16(o + 2 + 2 * n + 2 * i); }
    QD_HD const u8 *sub(u32 k) const { return o + 4 + 4 * n + 16 * size_t(k); }
    QD_HD i16 sub_hmlz(u32 k) const { return ldi16(sub(k)); }
    QD_HD i16 sub_type(u32 k) const { return ldi16(sub(k) + 2); }
    QD_HD u64 sub_hash(u32 k) const { return ld64(sub(k) + 8); }
};

struct QcgdCoef {
    cplx d, dn, dc, dnc;  // do_, do_not, do_conj, do_not_conj
};
QD_HD QcgdCoef coef_of(const qd_rule &r) {
    return {{r.do_re, r.do_im}, {r.do_not_re, r.do_not_im}, {r.do_conj_re, r.do_conj_im},
            {r.do_not_conj_re, r.do_not_conj_im}};
}

// Root hash of the .l / .r half of a node whose name tree starts at sub
// index b (operations::left / right, qcgd.hpp:194-208, with sub_node(node,
// type), 43-51: the int16 type is sign-extended to size_t).
QD_HD u64 split_left_hash(const Graph &g, u32 b) {
    const i16 t = g.sub_type(b);
    if (t >= 0) return g.sub_hash(b + 1);
    return hash_combine(g.sub_hash(b), u64(int64_t(QCGD_DOT_L)));
}
QD_HD u64 split_right_hash(const Graph &g, u32 b) {
    const i16 t = g.sub_type(b);
    if (t >= 0) return g.sub_hash(b + u32(t));
    return hash_combine(g.sub_hash(b), u64(int64_t(QCGD_DOT_R)));
}
// Root hash of merge(node a, node b), operations::merge, qcgd.hpp:181-192.
QD_HD u64 merge_hash(const Graph &g, u32 ba, u32 bb) {
    if (g.sub_type(ba) == QCGD_DOT_L && g.sub_type(bb) == QCGD_DOT_R && g.sub_hash(ba + 1) == g.sub_hash(bb + 1))
        return g.sub_hash(ba + 1);
    return hash_combine(g.sub_hash(ba), g.sub_hash(bb));
}
// Number of sub_nodes of those trees.
QD_HD u32 split_left_len(const Graph &g, u32 b, u32 m) {
    const i16 t = g.sub_type(b);
    return t >= 0 ? u32(t) - 1 : m + 1;
}
QD_HD u32 split_right_len(const Graph &g, u32 b, u32 m) {
    const i16 t = g.sub_type(b);
    return t >= 0 ? m - u32(t) : m + 1;
}
QD_HD bool merge_collapses(const Graph &g, u32 ba, u32 bb) {
    return g.sub_type(ba) == QCGD_DOT_L && g.sub_type(bb) == QCGD_DOT_R && g.sub_hash(ba + 1) == g.sub_hash(bb + 1);
}

// qcgd.hpp:122-146 over an encoded graph.
QD_HD u64 hash_graph(const u8 *obj) {
    Graph g(obj);
    u64 lh = 0, rh = 0, nh = 0;
    for (u32 i = 0; i < g.n; ++i) {
        if (g.L(i)) lh = hash_combine_small(lh, i);
        if (g.R(i)) rh = hash_combine_small(rh, i);
        nh = hash_combine(nh, g.sub_hash(g.name_begin(i)));
    }
    nh = hash_combine(nh, lh);
    return hash_combine(nh, rh);
}

// Fold over the node roots only (the part of hash_graph that erase_create /
// coin children share with their parent).
QD_HD u64 name_fold(const Graph &g) {
    u64 nh = 0;
    for (u32 i = 0; i < g.n; ++i) nh = hash_combine(nh, g.sub_hash(g.name_begin(i)));
    return nh;
}

// ---- erase_create / coin -------------------------------------------------
QD_HD bool flip_eligible(bool coin, bool l, bool r) { return coin ? (l != r) : (l == r); }

QD_HD u32 flip_num_child(const Graph &g, bool coin) {
    u32 nc = 1;
    for (u32 i = 0; i < g.n; ++i)
        if (flip_eligible(coin, g.L(i), g.R(i))) nc *= 2;  // u32 wraps like the reference's uint
    return nc;
}

// Child hash + magnitude: erase_create::populate_child (qcgd.hpp:487-510),
// coin::populate_child (562-584) followed by hash_graph; nf = name_fold.
template <class Coef, class Mag>
QD_HD u64 flip_child_hash(const Graph &g, bool coin, u64 nf, u32 child_id, const Coef &k, Mag &mag) {
    u32 cid = child_id;
    u64 lh = 0, rh = 0;
    for (u32 i = 0; i < g.n; ++i) {
        bool l = g.L(i), r = g.R(i);
        if (flip_eligible(coin, l, r)) {
            const bool conj = l;  // erase_create: sum/2 == l when l == r; coin: left
            if (cid & 1) {
                mag = cmul(mag, conj ? k.dc : k.d);
                l = !l;
                r = !r;
            } else {
                mag = cmul(mag, conj ? cneg(k.dnc) : k.dn);
            }
            cid >>= 1;
        }
        if (l) lh = hash_combine_small(lh, i);
        if (r) rh = hash_combine_small(rh, i);
    }
    return hash_combine(hash_combine(nf, lh), rh);
}

// populate_child_simple (qcgd.hpp:511-531, 585-604): the parent bytes with
// particle bits flipped.
QD_HD void flip_child_write(const u8 *parent, u32 size, bool coin, u32 child_id, u8 *out) {
    for (u32 i = 0; i < size; ++i) out[i] = parent[i];
    const u32 n = ld16(parent);
    u32 cid = child_id;
    for (u32 i = 0; i < n; ++i) {
        const bool l = parent[2 + i] != 0, r = parent[2 + n + i] != 0;
        if (flip_eligible(coin, l, r)) {
            if (cid & 1) {
                out[2 + i] = !l;
                out[2 + n + i] = !r;
            }
            cid >>= 1;
        }
    }
}

// ---- split_merge ---------------------------------------------------------
// operations::get_operations, qcgd.hpp:174-179
QD_HD void sm_ops(const Graph &g, u32 i, bool &split, bool &merge) {
    split = g.L(i) && g.R(i);
    merge = !split && i + 1 < g.n && g.L(i) && g.R(i + 1) && !g.L(i + 1);
}

struct SmHeader {
    bool first_split, last_merge;
};
QD_HD SmHeader sm_header(const Graph &g) {
    SmHeader h;
    h.first_split = g.L(0) && g.R(0);
    h.last_merge = false;
    if (!h.first_split && g.n > 1) h.last_merge = g.R(0) && g.L(g.n - 1) && !g.R(g.n - 1);
    return h;
}

// split_merge::get_num_child, qcgd.hpp:623-642
QD_HD u32 sm_num_child(const Graph &g) {
    const SmHeader h = sm_header(g);
    u32 nc = (h.first_split || h.last_merge) ? 2 : 1;
    for (int i = int(h.first_split); i < int(g.n) - int(h.last_merge); ++i) {
        bool s, m;
        sm_ops(g, u32(i), s, m);
        if (s || m) nc *= 2;
    }
    return nc;
}

// has_most_left_zero of the first node (qcgd.hpp:709-712): false only when
// node 0 is a pair whose left half has a positive hmlz_and_element.
QD_HD bool sm_first_hmlz(const Graph &g) {
    const u32 b = g.name_begin(0);
    return !(g.sub_type(b) >= 0 && g.sub_hmlz(b + 1) > 0);
}

// Per-parent table for split_merge children: for every node i, three u64
// root hashes {own, split-left or merge, split-right}.  A node holding both
// particles (a split) never merges, so its second slot holds the left half
// of the split; any other node's holds the merge with node (i+1) mod n
// (the wrap-around last merge of node n-1 included).
constexpr u32 kSmSlots = 3;
QD_HD u32 sm_slot(u32 i, int kind) { return kSmSlots * i + (kind == 0 ? 0u : kind == 2 ? 2u : 1u); }
QD_HD void sm_fill_table(const Graph &g, u64 *tab) {
    for (u32 i = 0; i < g.n; ++i) {
        const u32 b = g.name_begin(i);
        const bool split = g.L(i) && g.R(i);
        tab[kSmSlots * i + 0] = g.sub_hash(b);
        tab[kSmSlots * i + 1] = split ? split_left_hash(g, b) : merge_hash(g, b, g.name_begin(i + 1 < g.n ? i + 1 : 0));
        tab[kSmSlots * i + 2] = split ? split_right_hash(g, b) : 0;
    }
}
// First table slot of the parent at byte offset `begin`: a valid graph of n
// nodes spans at least 4 + 20 n bytes (every node names at least one
// 16-byte sub_node), so floor(3 begin / 20) leaves 3 n slots before the next
// parent's; the table holds 3 / 20 of the state's bytes in u64 slots.
QD_HD u64 sm_tab_base(u64 begin) { return begin * kSmSlots / 20; }

// Walks the child node sequence of split_merge::populate_child
// (qcgd.hpp:643-848): calls emit(kind, i, l, r) for every child node in
// order, kind 0 = copy of node i, 1 = left half of i, 2 = right half of i,
// 3 = merge of i with (i+1) mod n; multiplies mag with the reference's
// factors in the reference's order when do_mag.
// Coef / Mag may be the dual-number types of the adjoint (value plus
// derivatives), for which cmul / cneg are overloaded.
template <class Coef, class Mag, class Emit>
QD_HD void sm_walk(const Graph &g, bool first_hmlz, u32 child_id, const Coef &k, Mag &mag, bool do_mag,
                   Emit &&emit) {
    const u32 n = g.n;
    SmHeader h = sm_header(g);
    u32 cid = child_id;
    const bool fs = h.first_split && (cid & 1);
    if (fs) {
        if (do_mag) mag = cmul(mag, k.d);
        cid >>= 1;
    }
    bool lm = h.last_merge;
    if (lm) {
        if (cid & 1) {
            if (do_mag) mag = cmul(mag, k.dc);
        } else {
            lm = false;
            if (do_mag) mag = cmul(mag, cneg(k.dnc));
        }
        cid >>= 1;
    }
    bool overflow = false;
    if (fs) {
        if (first_hmlz) {
            emit(1, 0u, true, false);
            emit(2, 0u, false, true);
        } else {
            overflow = true;
            emit(2, 0u, false, true);
        }
    }
    if (lm) emit(3, n - 1, true, true);
    // A merge at node i implies no operation at node i+1 (it needs
    // left[i+1] false), so skipping i+1 after a merge, as the reference's
    // second pass does, skips no factor of its first pass: the factors come
    // out in the same node order.
    for (u32 i = u32(fs) + u32(lm); i < n - u32(lm); ++i) {
        bool s, m;
        sm_ops(g, i, s, m);
        if (s || m) {
            if (cid & 1) {
                if (s) {
                    if (do_mag) mag = cmul(mag, k.d);
                    emit(1, i, true, false);
                    emit(2, i, false, true);
                } else {
                    if (do_mag) mag = cmul(mag, k.dc);
                    emit(3, i, true, true);
                    cid >>= 1;
                    ++i;
                    continue;
                }
            } else {
                if (do_mag) mag = cmul(mag, s ? k.dn : cneg(k.dnc));
                emit(0, i, g.L(i), g.R(i));
            }
            cid >>= 1;
        } else {
            emit(0, i, g.L(i), g.R(i));
        }
    }
    if (overflow) emit(1, 0u, true, false);
}

// Child hash + magnitude from the per-parent table.
template <class Coef, class Mag>
QD_HD u64 sm_child_hash(const Graph &g, bool first_hmlz, const u64 *tab, u32 child_id, const Coef &k, Mag &mag) {
    u64 nh = 0, lh = 0, rh = 0;
    u32 c = 0;
    sm_walk(g, first_hmlz, child_id, k, mag, true, [&](int kind, u32 i, bool l, bool r) {
        nh = hash_combine(nh, tab[sm_slot(i, kind)]);
        if (l) lh = hash_combine_small(lh, c);
        if (r) rh = hash_combine_small(rh, c);
        ++c;
    });
    return hash_combine(hash_combine(nh, lh), rh);
}

// Child size in bytes: 4 + 4 n_child + 16 * (sub_nodes).
QD_HD u32 sm_child_size(const Graph &g, bool first_hmlz, u32 child_id) {
    u32 nodes = 0, subs = 0;
    cplx dummy{0, 0};
    QcgdCoef k{};
    sm_walk(g, first_hmlz, child_id, k, dummy, false, [&](int kind, u32 i, bool, bool) {
        const u32 b = g.name_begin(i);
        const u32 m = g.name_begin(i + 1) - b;
        ++nodes;
        if (kind == 0) subs += m;
        else if (kind == 1) subs += split_left_len(g, b, m);
        else if (kind == 2) subs += split_right_len(g, b, m);
        else {
            const u32 j = i + 1 < g.n ? i + 1 : 0;
            const u32 bb = g.name_begin(j);
            const u32 mb = g.name_begin(j + 1) - bb;
            subs += merge_collapses(g, b, bb) ? m - 1 : 1 + m + mb;
        }
    });
    return 4 + 4 * nodes + 16 * subs;
}

// Writes one sub_node with zero padding.
QD_HD void put_sub(u8 *d, i16 hmlz, i16 type, u64 hash) {
    st16(d, u16(hmlz));
    st16(d + 2, u16(type));
    for (int i = 4; i < 8; ++i) d[i] = 0;
    for (int i = 0; i < 8; ++i) d[8 + i] = u8(hash >> (8 * i));
}
QD_HD void copy_sub(u8 *d, const u8 *s) {
    for (int i = 0; i < 16; ++i) d[i] = s[i];
}
QD_HD void copy_subs(u8 *d, const Graph &g, u32 b, u32 e) {
    for (u32 k = b; k < e; ++k) copy_sub(d + 16 * (k - b), g.sub(k));
}

// split_merge::populate_child_simple (qcgd.hpp:850-1035): writes the child.
// Returns its size.
QD_HD u32 sm_child_write(const Graph &g, bool first_hmlz, u32 child_id, u8 *out) {
    // pass 1: node count
    u32 nodes = 0;
    {
        cplx dm{0, 0};
        QcgdCoef k{};
        sm_walk(g, first_hmlz, child_id, k, dm, false, [&](int, u32, bool, bool) { ++nodes; });
    }
    st16(out, u16(nodes));
    u8 *Lc = out + 2, *Rc = out + 2 + nodes, *nb = out + 2 + 2 * nodes, *subs = out + 4 + 4 * nodes;
    u32 c = 0, off = 0;
    cplx dm{0, 0};
    QcgdCoef k{};
    sm_walk(g, first_hmlz, child_id, k, dm, false, [&](int kind, u32 i, bool l, bool r) {
        Lc[c] = l;
        Rc[c] = r;
        st16(nb + 2 * c, u16(off));
        const u32 b = g.name_begin(i), e = g.name_begin(i + 1), m = e - b;
        u8 *d = subs + 16 * size_t(off);
        if (kind == 0) {
            copy_subs(d, g, b, e);
            off += m;
        } else if (kind == 1 || kind == 2) {
            const i16 t = g.sub_type(b);
            if (t >= 0) {
                // pair: copy the half (left: [b+1, b+t), right: [b+t, e))
                const u32 hb = kind == 1 ? b + 1 : b + u32(t);
                const u32 he = kind == 1 ? b + u32(t) : e;
                copy_subs(d, g, hb, he);
                off += he - hb;
            } else {
                const i16 type = kind == 1 ? QCGD_DOT_L : QCGD_DOT_R;
                const i16 hm = (type == QCGD_DOT_L && g.sub_hmlz(b) < 0) ? i16(-1) : i16(1);
                put_sub(d, hm, type, kind == 1 ? split_left_hash(g, b) : split_right_hash(g, b));
                copy_subs(d + 16, g, b, e);
                off += 1 + m;
            }
        } else {
            const u32 j = i + 1 < g.n ? i + 1 : 0;
            const u32 bb = g.name_begin(j), eb = g.name_begin(j + 1), mb = eb - bb;
            if (merge_collapses(g, b, bb)) {
                copy_subs(d, g, b + 1, e);
                off += m - 1;
            } else {
                const i16 hm = (g.sub_hmlz(b) < 0 || g.sub_hmlz(bb) < 0) ? i16(-1) : i16(1);
                put_sub(d, hm, i16(m + 1), hash_combine(g.sub_hash(b), g.sub_hash(bb)));
                copy_subs(d + 16, g, b, e);
                copy_subs(d + 16 * (1 + m), g, bb, eb);
                off += 1 + m + mb;
            }
        }
        ++c;
    });
    st16(nb + 2 * nodes, u16(off));
    return 4 + 4 * nodes + 16 * off;
}

// ---- quantum_computer ----------------------------------------------------
// hadamard::populate_child, quantum_computer.hpp:40-49
QD_HD u64 hadamard_child_hash(const u8 *p, u32 size, u32 bit, double coef, u32 child_id, cplx &mag) {
    const bool b = p[bit] != 0;
    mag = cscale(mag, (b && child_id) ? -coef : coef);
    return hash_bytes_flip(p, size, bit, u8(!child_id));
}
QD_HD void hadamard_child_write(const u8 *p, u32 size, u32 bit, u32 child_id, u8 *out) {
    for (u32 i = 0; i < size; ++i) out[i] = p[i];
    out[bit] ^= u8(!child_id);
}

// ---------------------------------------------------------------------------
// generic dispatch used by host calls and by the slow (byte-level) kernels
// ---------------------------------------------------------------------------
QD_HD bool rule_known(int kind) {
    return kind == QD_RULE_HADAMARD || kind == QD_RULE_QCGD_ERASE_CREATE || kind == QD_RULE_QCGD_COIN ||
           kind == QD_RULE_QCGD_SPLIT_MERGE;
}

QD_HD u32 rule_num_child(const qd_rule &r, const u8 *p, u32 size, u32 *max_child_size) {
    switch (r.kind) {
    case QD_RULE_HADAMARD:
        *max_child_size = size;
        return 2;
    case QD_RULE_QCGD_ERASE_CREATE:
        *max_child_size = size;
        return flip_num_child(Graph(p), false);
    case QD_RULE_QCGD_COIN:
        *max_child_size = size;
        return flip_num_child(Graph(p), true);
    case QD_RULE_QCGD_SPLIT_MERGE:
        *max_child_size = 4 * size;
        return sm_num_child(Graph(p));
    }
    *max_child_size = 0;
    return 0;
}

QD_HD u64 rule_hash_object(const qd_rule &r, const u8 *obj, u32 size) {
    if (r.kind == QD_RULE_HADAMARD) return hash_bytes(obj, size);
    return hash_graph(obj);
}

// Child size without writing it.
QD_HD u32 rule_child_size(const qd_rule &r, const u8 *p, u32 size, u32 child_id) {
    if (r.kind == QD_RULE_QCGD_SPLIT_MERGE) {
        Graph g(p);
        return sm_child_size(g, sm_first_hmlz(g), child_id);
    }
    return size;
}

// populate_child_simple: writes the child, returns its size.
QD_HD u32 rule_child_write(const qd_rule &r, const u8 *p, u32 size, u32 child_id, u8 *out) {
    switch (r.kind) {
    case QD_RULE_HADAMARD:
        hadamard_child_write(p, size, r.bit, child_id, out);
        return size;
    case QD_RULE_QCGD_ERASE_CREATE:
    case QD_RULE_QCGD_COIN:
        flip_child_write(p, size, r.kind == QD_RULE_QCGD_COIN, child_id, out);
        return size;
    case QD_RULE_QCGD_SPLIT_MERGE: {
        Graph g(p);
        return sm_child_write(g, sm_first_hmlz(g), child_id, out);
    }
    }
    return 0;
}

// Child magnitude only (populate_child's mag output).
QD_HD cplx rule_child_mag(const qd_rule &r, const u8 *p, u32 size, u32 child_id, cplx mag) {
    const QcgdCoef k = coef_of(r);
    switch (r.kind) {
    case QD_RULE_HADAMARD:
        hadamard_child_hash(p, size, r.bit, r.coef, child_id, mag);
        return mag;
    case QD_RULE_QCGD_ERASE_CREATE:
    case QD_RULE_QCGD_COIN: {
        Graph g(p);
        flip_child_hash(g, r.kind == QD_RULE_QCGD_COIN, 0, child_id, k, mag);
        return mag;
    }
    case QD_RULE_QCGD_SPLIT_MERGE: {
        Graph g(p);
        sm_walk(g, true, child_id, k, mag, true, [](int, u32, bool, bool) {});
        return mag;
    }
    }
    return mag;
}

// ---- modifiers -----------------------------------------------------------
QD_HD bool modifier_known(int kind) { return kind >= QD_MOD_CNOT && kind <= QD_MOD_QCGD_REVERSED_STEP; }

QD_HD void modifier_apply(const qd_modifier &m, u8 *obj, u32 size, cplx &mag) {
    switch (m.kind) {
    case QD_MOD_CNOT:  // quantum_computer.hpp:25-29
        obj[m.bit] ^= obj[m.control_bit];
        break;
    case QD_MOD_XGATE:  // 52-56
        obj[m.bit] = !obj[m.bit];
        break;
    case QD_MOD_YGATE:  // 58-66
        mag = cmul(mag, cplx{0.0, 1.0});
        if (obj[m.bit]) mag = cscale(mag, -1.0);
        obj[m.bit] = !obj[m.bit];
        break;
    case QD_MOD_ZGATE:  // 68-75: multiplies by -1 when set, and flips the bit
        if (obj[m.bit]) mag = cscale(mag, -1.0);
        obj[m.bit] = !obj[m.bit];
        break;
    case QD_MOD_QCGD_STEP:             // qcgd.hpp:443-449: left[i] <- left[i+1],
    case QD_MOD_QCGD_REVERSED_STEP: {  // right[i] <- right[i-1]; reversed: inverse
        const u32 n = ld16(obj);
        if (n == 0) break;
        u8 *L = obj + 2, *R = obj + 2 + n;
        const bool fwd = m.kind == QD_MOD_QCGD_STEP;
        u8 *A = fwd ? L : R;  // rotated towards lower indices
        u8 *B = fwd ? R : L;  // rotated towards higher indices
        const u8 a0 = A[0];
        for (u32 i = 0; i + 1 < n; ++i) A[i] = A[i + 1];
        A[n - 1] = a0;
        const u8 bl = B[n - 1];
        for (u32 i = n - 1; i > 0; --i) B[i] = B[i - 1];
        B[0] = bl;
        break;
    }
    }
    (void)size;
}

// Magnitude factor a modifier applied, read from the object AFTER the
// modifier (Y and Z flip their bit, so the bit before is the negation of the
// bit after): the adjoint multiplies cotangents by its conjugate.
QD_HD cplx modifier_factor_after(const qd_modifier &m, const u8 *obj) {
    switch (m.kind) {
    case QD_MOD_YGATE:  // i, times -1 when the bit was set
        return obj[m.bit] ? cplx{0.0, 1.0} : cplx{-0.0, -1.0};
    case QD_MOD_ZGATE:  // -1 when the bit was set
        return obj[m.bit] ? cplx{1.0, 0.0} : cplx{-1.0, 0.0};
    default:
        return cplx{1.0, 0.0};
    }
}

// ---- random keys of probabilistic truncation -----------------------------
// The reference keeps the objects with the smallest rng() / |m|^2
// (quids.hpp:594-608, 829-845, 884-895), rng() from per-thread generators
// seeded with rand().  Here u is drawn from Philox2x64-10 (Salmon et al.,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011) with counter
// (object hash, 0) and key `seed`, so the draw depends only on the object
// and the seed.  Keeping the smallest u / |m|^2 is keeping the largest
// |m|^2 / u, which the top-K select handles like |m|^2.
QD_HD u64 mulhi64(u64 a, u64 b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return u64((unsigned __int128)a * b >> 64);
#endif
}
QD_HD u64 philox2x64_10(u64 c0, u64 c1, u64 key) {
    for (int r = 0; r < 10; ++r) {
        const u64 hi = mulhi64(0xD2B74407B1CE6E93ull, c0);
        const u64 lo = 0xD2B74407B1CE6E93ull * c0;
        c0 = hi ^ key ^ c1;
        c1 = lo;
        key += 0x9E3779B97F4A7C15ull;
    }
    return c0;
}
// uniform in (0, 1] with 53 random bits
QD_HD double trunc_uniform(u64 hash, u64 seed) {
    return double((philox2x64_10(hash, 0, seed) >> 11) + 1) * 0x1p-53;
}
// The high word of the select key: the bits of |m|^2, or of |m|^2 / u when
// truncation is probabilistic (both non-negative doubles, so their bit
// patterns order like their values).
struct SelKey {
    u32 prob;
    u32 pad;
    u64 seed;
};
QD_HD u64 sel_hi(double re, double im, u64 hash, const SelKey &k) {
    double v = re * re + im * im;
    if (k.prob) v = v / trunc_uniform(hash, k.seed);
    u64 b;
    memcpy(&b, &v, 8);
    return b;
}

// ---- dual numbers for the adjoint ----------------------------------------
// A complex value with its derivatives along up to kMaxDir real parameters
// of a rule (theta, phi, xi).  The rule walks above multiply these in the
// same factor order as the plain magnitudes.
constexpr int kMaxDir = 3;
struct DualC {
    cplx v;
    cplx t[kMaxDir];
};
QD_HD cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
QD_HD DualC cneg(const DualC &a) {
    DualC r;
    r.v = cneg(a.v);
    for (int i = 0; i < kMaxDir; ++i) r.t[i] = cneg(a.t[i]);
    return r;
}
QD_HD DualC cmul(const DualC &a, const DualC &b) {
    DualC r;
    r.v = cmul(a.v, b.v);
    for (int i = 0; i < kMaxDir; ++i) r.t[i] = cadd(cmul(a.t[i], b.v), cmul(a.v, b.t[i]));
    return r;
}
struct DualCoef {
    DualC d, dn, dc, dnc;  // do_, do_not, do_conj, do_not_conj with derivatives
};
// dr[i] holds, in its coefficient fields, the derivatives of r's
// coefficients along parameter i (i < ndir); further directions are zero.
QD_HD DualCoef dual_coef_of(const qd_rule &r, const qd_rule *dr, int ndir) {
    DualCoef k;
    const QcgdCoef v = coef_of(r);
    k.d.v = v.d;
    k.dn.v = v.dn;
    k.dc.v = v.dc;
    k.dnc.v = v.dnc;
    for (int i = 0; i < kMaxDir; ++i) {
        const QcgdCoef t = i < ndir ? coef_of(dr[i]) : QcgdCoef{};
        k.d.t[i] = t.d;
        k.dn.t[i] = t.dn;
        k.dc.t[i] = t.dc;
        k.dnc.t[i] = t.dnc;
    }
    return k;
}
QD_HD DualC dual_one() {
    DualC r{};
    r.v = cplx{1.0, 0.0};
    return r;
}

}  // namespace qd

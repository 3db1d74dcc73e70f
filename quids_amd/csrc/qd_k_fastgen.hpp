#pragma once
// Fast path of prep + generate for the QCGD rules (graphs of at most 64
// nodes): per-parent descriptors with bit masks, and child generation in
// work items of up to 8 siblings that share the walk over their common
// leading decisions.
//
// Decisions of a parent (src/rules/qcgd.hpp): split_merge consumes one
// child_id bit for the wrap-around last merge when it is possible
// (qcgd.hpp:648-668), then one per node holding a split or merge operation,
// in node order; the first node's split (qcgd.hpp:649, 655-659) is that
// node's decision.  erase_create / coin consume one bit per eligible node
// in node order (qcgd.hpp:494-509, 569-583).  Children that agree on their
// first s bits therefore agree on everything the walk computes before the
// s-th decision: the magnitude product so far, the folded hashes and the
// child node count.  A work item fixes the low bits j0 of child_id and
// enumerates its top t = min(k, 3) bits; the shared walk runs once.
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

constexpr u32 kItemBits = 3;  // up to 8 children per work item

struct __align__(16) PDesc {
    u64 lmask, rmask;  // particle bits
    u64 dmask;         // nodes that carry a decision
    u64 nf;            // erase_create / coin: fold of the node root hashes, times kMul
    u32 nc;            // num_child, 2^k wrapped to 32 bits like the reference's uint
    u8 n, k, flags, pad;
};
enum : u8 { PD_FS = 1, PD_LM = 2, PD_HMLZ = 4 };

// Reported through the error word: a parent with more than 64 nodes (the
// step then runs on the generic per-byte kernels).
constexpr u32 ERR_WIDE = 8u;

// Descriptor of parent p, computed by a segment of SEG lanes (one lane per
// node: SEG = 64, or 32 / 16 when two / four parents share a wavefront; the
// segment's lanes are active together).  have: the segment
// has a parent.  Ballots and shuffles stay within the segment.
// NC_ONLY: the child count alone (nc; D, wi and tab untouched).
template <int KIND, int SEG, bool NC_ONLY = false>
__device__ inline void prep_parent(const Parents &P, u64 p, bool have, const u64 *keep, PDesc *D, u32 *nc, u32 *wi,
                                   u64 *tab, u32 &local_max, u32 *err) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u32 sl = lane & (SEG - 1);      // node of this lane
    const u32 sbase = lane & ~(SEG - 1);  // first lane of the segment
    const u64 smask = SEG == kWave ? ~0ull : (((1ull << SEG) - 1) << sbase);
    auto sballot = [&](bool x) -> u64 { return (u64(__ballot(x)) & smask) >> sbase; };
    if (!have) return;
    const u8 *o = P.obj + P.begin[p];
    const u32 sz = P.size[p];
    PDesc d{};
    bool ok = !(keep && !((keep[p >> 6] >> (p & 63)) & 1));
    // qcgd_valid, one node per lane
    const u32 n = (ok && sz >= 4) ? ld16(o) : 0;
    if (ok) {
        bool valid = n > 0 && 4 + 4 * u64(n) <= sz;
        if (valid) {
            const u8 *nb = o + 2 + 2 * n;
            const u32 m = ld16(nb + 2 * n);
            valid = ld16(nb) == 0 && 4 + 4 * u64(n) + 16 * u64(m) == sz;
            if (valid && n <= SEG) {
                bool bad = false;
                if (sl < n) {
                    const u32 b = ld16(nb + 2 * sl), e = ld16(nb + 2 * sl + 2);
                    if (e <= b || e > m) bad = true;
                    else {
                        const u32 mi = e - b;
                        const i16 t = ldi16(o + 4 + 4 * n + 16 * size_t(b) + 2);
                        if (t >= 0) bad = t < 2 || u32(t) >= mi;
                        else if (t == QCGD_DOT_L || t == QCGD_DOT_R) bad = mi < 2;
                        else bad = t != QCGD_ELEMENT;
                    }
                }
                valid = sballot(bad) == 0;
            }
        }
        if (!valid) {
            if (sl == 0) atomicOr(err, ERR_OBJECT);
            ok = false;
        }
    }
    if (ok && n > 64) {
        if (sl == 0) atomicOr(err, ERR_WIDE);
        ok = false;
    }
    if (!ok) {
        if (sl == 0) {
            nc[p] = 0;
            if (!NC_ONLY) {
                wi[p] = 0;
                D[p] = d;
            }
        }
        return;
    }
    Graph g(o);
    const bool in = sl < n;
    const bool Li = in && g.L(sl), Ri = in && g.R(sl);
    const u64 lm = sballot(Li), rm = sballot(Ri);
    d.lmask = lm;
    d.rmask = rm;
    d.n = u8(n);
    u32 k = 0;
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
        const bool Ln = sl + 1 < n && ((lm >> (sl + 1)) & 1);
        const bool Rn = sl + 1 < n && ((rm >> (sl + 1)) & 1);
        const bool split = Li && Ri;
        const bool merge = !split && sl + 1 < n && Li && Rn && !Ln;
        d.dmask = sballot(in && (split || merge));
        const bool fs = (lm & 1) && (rm & 1);
        const bool lmg = !fs && n > 1 && (rm & 1) && ((lm >> (n - 1)) & 1) && !((rm >> (n - 1)) & 1);
        d.flags = u8((fs ? PD_FS : 0) | (lmg ? PD_LM : 0));
        if (!NC_ONLY) {
            if (sl == 0 && sm_first_hmlz(g)) d.flags |= PD_HMLZ;
            d.flags = u8(__shfl(int(d.flags), int(sbase), kWave));
        }
        k = u32(__popcll(d.dmask)) + (lmg ? 1 : 0);
        if (!NC_ONLY && in && tab) {  // the slots of sm_fill_table (tab null: filled by a later launch)
            u64 *t = tab + sm_tab_base(P.begin[p]) + kSmSlots * sl;
            const u32 b = g.name_begin(sl);
            t[0] = g.sub_hash(b);
            t[1] = split ? split_left_hash(g, b) : merge_hash(g, b, g.name_begin(sl + 1 < n ? sl + 1 : 0));
            t[2] = split ? split_right_hash(g, b) : 0;
        }
        local_max = max(local_max, 4 * sz);
    } else {
        const bool coin = KIND == QD_RULE_QCGD_COIN;
        d.dmask = sballot(in && (coin ? (Li != Ri) : (Li == Ri)));
        k = u32(__popcll(d.dmask));
        if (!NC_ONLY) {
            const u64 hr = in ? g.sub_hash(g.name_begin(sl)) : 0;
            u64 nf = 0;
            for (u32 i = 0; i < n; ++i) nf = hash_combine(nf, __shfl(hr, int(sbase + i), kWave));
            d.nf = nf * kMul;  // pre-multiplied, as the walks' folds carry it
        }
        local_max = max(local_max, sz);
    }
    d.k = u8(k);
    d.nc = k >= 32 ? 0u : (1u << k);
    if (sl == 0) {
        nc[p] = d.nc;
        if (!NC_ONLY) {
            const u32 t = min(k, kItemBits);
            wi[p] = d.nc >> t;
            D[p] = d;
        }
    }
}

// Four parents per wavefront: in 16-lane segments when none has more than
// 16 nodes, two by two in 32-lane segments when none has more than 32, else
// one after the other on the whole wavefront.
template <int KIND>
__global__ __launch_bounds__(256) void k_prep_fast(Parents P, const u64 *keep, PDesc *D, u32 *nc, u32 *wi, u64 *tab,
                                                   u32 *maxsz, u32 *err) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u64 nwaves = u64(gridDim.x) * (blockDim.x / kWave);
    u32 local_max = 0;
    for (u64 p0 = 4 * (u64(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave); p0 < P.n; p0 += 4 * nwaves) {
        u32 nq = 0;  // node count of parent p0 + lane (lanes 0..3), as far as its bytes tell
        if (lane < 4 && p0 + lane < P.n && P.size[p0 + lane] >= 4) nq = ld16(P.obj + P.begin[p0 + lane]);
        if (__ballot(nq > 16) == 0) {
            const u64 p = p0 + lane / 16;
            prep_parent<KIND, 16>(P, p, p < P.n, keep, D, nc, wi, tab, local_max, err);
        } else if (__ballot(nq > 32) == 0) {
            for (u32 h = 0; h < 2; ++h) {
                const u64 p = p0 + 2 * h + lane / 32;
                prep_parent<KIND, 32>(P, p, p < P.n, keep, D, nc, wi, tab, local_max, err);
            }
        } else {
            for (u32 q = 0; q < 4; ++q)
                prep_parent<KIND, kWave>(P, p0 + q, p0 + q < P.n, keep, D, nc, wi, tab, local_max, err);
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) local_max = max(local_max, (u32)__shfl_xor(int(local_max), off, kWave));
    if (lane == 0) raise_max(maxsz, local_max);
}

// The child counts alone, in k_prep_fast's segments (the batched step's
// planning pass): the same checks and counts as the step's prep.  A graph
// of more than 64 nodes raises ERR_WIDE and counts 0; the caller then
// counts with k_num_child.
template <int KIND>
__global__ __launch_bounds__(256) void k_num_child_fast(Parents P, u32 *nc, u32 *err) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u64 nwaves = u64(gridDim.x) * (blockDim.x / kWave);
    u32 local_max = 0;
    for (u64 p0 = 4 * (u64(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave); p0 < P.n; p0 += 4 * nwaves) {
        u32 nq = 0;
        if (lane < 4 && p0 + lane < P.n && P.size[p0 + lane] >= 4) nq = ld16(P.obj + P.begin[p0 + lane]);
        if (__ballot(nq > 16) == 0) {
            const u64 p = p0 + lane / 16;
            prep_parent<KIND, 16, true>(P, p, p < P.n, nullptr, nullptr, nc, nullptr, nullptr, local_max, err);
        } else if (__ballot(nq > 32) == 0) {
            for (u32 h = 0; h < 2; ++h) {
                const u64 p = p0 + 2 * h + lane / 32;
                prep_parent<KIND, 32, true>(P, p, p < P.n, nullptr, nullptr, nc, nullptr, nullptr, local_max, err);
            }
        } else {
            for (u32 q = 0; q < 4; ++q)
                prep_parent<KIND, kWave, true>(P, p0 + q, p0 + q < P.n, nullptr, nullptr, nc, nullptr, nullptr,
                                               local_max, err);
        }
    }
}

// QCGD step / reversed_step (qcgd.hpp:443-457) on a segment of SEG lanes,
// one lane per node (SEG = 64, or 32 / 16 when two / four graphs share a
// wavefront; the segment's lanes are active together): left[i] <-
// left[i+1], right[i] <- right[i-1] (cyclic), or the inverse, by shuffles
// within the segment.  The object is validated as in prep; a graph wider
// than the segment is rotated by one lane.  Magnitudes are unchanged.
template <int SEG>
__device__ inline void step_object(u8 *obj, const u64 *begin, const u32 *size, u64 p, bool have, const qd_modifier &m,
                                   u32 *err) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u32 sl = lane & (SEG - 1);
    const u32 sbase = lane & ~(SEG - 1);
    const u64 smask = SEG == kWave ? ~0ull : (((1ull << SEG) - 1) << sbase);
    if (!have) return;
    u8 *o = obj + begin[p];
    const u32 sz = size[p];
    const u32 n = sz >= 4 ? ld16(o) : 0;
    bool valid = n > 0 && 4 + 4 * u64(n) <= sz;
    if (valid) {
        const u8 *nb = o + 2 + 2 * n;
        const u32 mm = ld16(nb + 2 * n);
        valid = ld16(nb) == 0 && 4 + 4 * u64(n) + 16 * u64(mm) == sz;
        if (valid && n <= SEG) {
            bool bad = false;
            if (sl < n) {
                const u32 b = ld16(nb + 2 * sl), e = ld16(nb + 2 * sl + 2);
                if (e <= b || e > mm) bad = true;
                else {
                    const u32 mi = e - b;
                    const i16 t = ldi16(o + 4 + 4 * n + 16 * size_t(b) + 2);
                    if (t >= 0) bad = t < 2 || u32(t) >= mi;
                    else if (t == QCGD_DOT_L || t == QCGD_DOT_R) bad = mi < 2;
                    else bad = t != QCGD_ELEMENT;
                }
            }
            valid = (u64(__ballot(bad)) & smask) == 0;
        } else if (valid) {
            valid = qcgd_valid(o, sz);  // wider than the segment: the scalar check
        }
    }
    if (!valid) {
        if (sl == 0) atomicOr(err, ERR_OBJECT);
        return;
    }
    if (n > SEG) {
        if (sl == 0) {
            cplx c{1.0, 0.0};
            modifier_apply(m, o, sz, c);
        }
        return;
    }
    const bool in = sl < n;
    const int l = in ? o[2 + sl] : 0, r = in ? o[2 + n + sl] : 0;
    const int up = int(sbase + (sl + 1 >= n ? 0 : sl + 1)), dn = int(sbase + (sl == 0 ? n - 1 : sl - 1));
    const bool fwd = m.kind == QD_MOD_QCGD_STEP;
    const int nl = __shfl(l, fwd ? up : dn, kWave), nr = __shfl(r, fwd ? dn : up, kWave);
    if (in) {
        o[2 + sl] = u8(nl);
        o[2 + n + sl] = u8(nr);
    }
}

// Four graphs per wavefront: in 16-lane segments when none has more than 16
// nodes, two by two in 32-lane segments when none has more than 32, else one
// after the other on the whole wavefront.
__global__ __launch_bounds__(256) void k_step_modifier(u8 *obj, const u64 *begin, const u32 *size, u64 N, qd_modifier m,
                                                       u32 *err) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u64 nwaves = u64(gridDim.x) * (blockDim.x / kWave);
    for (u64 p0 = 4 * (u64(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave); p0 < N; p0 += 4 * nwaves) {
        u32 nq = 0;
        if (lane < 4 && p0 + lane < N && size[p0 + lane] >= 4) nq = ld16(obj + begin[p0 + lane]);
        if (__ballot(nq > 16) == 0) {
            const u64 p = p0 + lane / 16;
            step_object<16>(obj, begin, size, p, p < N, m, err);
        } else if (__ballot(nq > 32) == 0) {
            for (u32 h = 0; h < 2; ++h) {
                const u64 p = p0 + 2 * h + lane / 32;
                step_object<32>(obj, begin, size, p, p < N, m, err);
            }
        } else {
            for (u32 q = 0; q < 4; ++q) step_object<kWave>(obj, begin, size, p0 + q, p0 + q < N, m, err);
        }
    }
}

// Walk state of one child (or of the shared prefix of a work item).  M is
// the magnitude type: cplx, or DualC (with derivatives, for the adjoint).
// The three hash folds are carried pre-multiplied (qd_rules.hpp): nh * kMul,
// lh * kMul^2, rh * kMul^2.
template <typename M>
struct WalkT {
    u64 nh, lh, rh;
    M mag;
    u32 i, end, c, d;
    bool started, overflow;
};
using Walk = WalkT<cplx>;

// split_merge child walk (populate_child, qcgd.hpp:643-848) over the
// parent's masks and root-hash table.  Consumes decisions d while d <
// dlimit, the bit of decision d being bit (d - d0) of `bits` where d0 is
// w.d on entry (`bits` a u32, or a u64 for a walk over more than 32
// decisions at once); returns when the next decision is not available or
// the walk is complete.
template <typename W, typename KC, typename BT>
__device__ inline void sm_run(W &w, const PDesc &pd, const u64 *tab, BT bits, u32 dlimit, const KC &k) {
    const u32 n = pd.n;
    const u32 d0 = w.d;
    auto emit = [&](u64 root, bool l, bool r) {
        w.nh = step_pm(w.nh, root);
        if (l) w.lh = small_step_pm(w.lh, w.c);
        if (r) w.rh = small_step_pm(w.rh, w.c);
        ++w.c;
    };
    if (!w.started) {
        if (pd.flags & PD_LM) {
            if (w.d >= dlimit) return;
            const bool take = (bits >> (w.d++ - d0)) & 1;
            if (take) {
                w.mag = cmul(w.mag, k.dc);
                emit(tab[kSmSlots * (n - 1) + 1], true, true);
                w.i = 1;
                w.end = n - 1;
            } else {
                w.mag = cmul(w.mag, cneg(k.dnc));
                w.i = 0;
                w.end = n;
            }
        } else if (pd.flags & PD_FS) {
            if (w.d >= dlimit) return;
            const bool take = (bits >> (w.d++ - d0)) & 1;
            if (take) {
                w.mag = cmul(w.mag, k.d);
                if (pd.flags & PD_HMLZ) {
                    emit(tab[1], true, false);
                    emit(tab[2], false, true);
                } else {
                    w.overflow = true;
                    emit(tab[2], false, true);
                }
            } else {
                w.mag = cmul(w.mag, k.dn);
                emit(tab[0], true, true);
            }
            w.i = 1;
            w.end = n;
        } else {
            w.i = 0;
            w.end = n;
        }
        w.started = true;
    }
    while (w.i < w.end) {
        const u32 i = w.i;
        const bool l = (pd.lmask >> i) & 1, r = (pd.rmask >> i) & 1;
        if ((pd.dmask >> i) & 1) {
            if (w.d >= dlimit) return;
            const bool take = (bits >> (w.d++ - d0)) & 1;
            if (l && r) {  // split
                if (take) {
                    w.mag = cmul(w.mag, k.d);
                    emit(tab[kSmSlots * i + 1], true, false);
                    emit(tab[kSmSlots * i + 2], false, true);
                } else {
                    w.mag = cmul(w.mag, k.dn);
                    emit(tab[kSmSlots * i], l, r);
                }
            } else {  // merge with node i+1 (which carries no decision)
                if (take) {
                    w.mag = cmul(w.mag, k.dc);
                    emit(tab[kSmSlots * i + 1], true, true);
                    ++w.i;
                } else {
                    w.mag = cmul(w.mag, cneg(k.dnc));
                    emit(tab[kSmSlots * i], l, r);
                }
            }
        } else {
            emit(tab[kSmSlots * i], l, r);
        }
        ++w.i;
    }
}

template <typename W>
__device__ inline u64 sm_finish(W &w, const u64 *tab) {
    if (w.overflow) {  // left half of the first node goes last (qcgd.hpp:838-845)
        w.nh = step_pm(w.nh, tab[1]);
        w.lh = small_step_pm(w.lh, w.c);
        ++w.c;
    }
    return finish_pm(w.nh, w.lh, w.rh);
}

// erase_create / coin child walk (qcgd.hpp:487-510, 562-584).
template <bool COIN, typename W, typename KC, typename BT>
__device__ inline void flip_run(W &w, const PDesc &pd, BT bits, u32 dlimit, const KC &k) {
    const u32 n = pd.n;
    const u32 d0 = w.d;
    while (w.i < n) {
        const u32 i = w.i;
        bool l = (pd.lmask >> i) & 1, r = (pd.rmask >> i) & 1;
        if ((pd.dmask >> i) & 1) {
            if (w.d >= dlimit) return;
            const bool take = (bits >> (w.d++ - d0)) & 1;
            if (take) {
                w.mag = cmul(w.mag, l ? k.dc : k.d);
                l = !l;
                r = !r;
            } else {
                w.mag = cmul(w.mag, l ? cneg(k.dnc) : k.dn);
            }
        }
        if (l) w.lh = small_step_pm(w.lh, i);
        if (r) w.rh = small_step_pm(w.rh, i);
        ++w.i;
    }
}

// The four factors of a decision in a table (LDS) indexed by
// (taken << 1) | x: x = 0 -> do_not / do_, x = 1 -> -do_not_conj / do_conj.
// A flip-rule decision node has x = its left bit; a split_merge operation
// x = 1 for a merge, 0 for a split (qcgd.hpp:466-471, 541-546, 614-619).
// One 16-byte LDS read per decision in place of a select among four
// uniform complex constants (about 30 instructions per decision).
__device__ inline void fill_factor_table(const QcgdCoef &k, cplx *ct) {
    ct[0] = k.dn;
    ct[1] = cneg(k.dnc);
    ct[2] = k.d;
    ct[3] = k.dc;
}

// One split_merge child walked on its own, with all its decision bits (bit
// d: decision d): sm_run and sm_finish with the per-node cases folded into
// selects, so that lanes holding different children run one path: a
// decision node multiplies the magnitude by the factor its case and its
// decision pick; every node emits its first root (its own name tree, or,
// taken, the split's left half or the merge), a taken split also its right
// half, and a taken merge skips the next node.  The same products, roots and
// folds in the same order as sm_run.
template <typename M>
__device__ inline u64 sm_child_m(const PDesc &pd, const u64 *tab, u64 bits, const cplx *ct, cplx &mag) {
    const u32 n = pd.n;
    const M lmask = M(pd.lmask), rmask = M(pd.rmask), dmask = M(pd.dmask);
    u64 nh = 0, lh = 0, rh = 0;
    u32 c = 0, d = 0, i = 0, end = n;
    bool overflow = false;
    auto emit = [&](u64 root, bool l, bool r) {
        nh = step_pm(nh, root);
        if (l) lh = small_step_pm(lh, c);
        if (r) rh = small_step_pm(rh, c);
        ++c;
    };
    if (pd.flags & PD_LM) {  // the last node merges with the first (qcgd.hpp:647-668)
        const bool take = bits & 1;
        d = 1;
        mag = cmul(mag, ct[take ? 3u : 1u]);  // a merge: do_conj / -do_not_conj
        if (take) {
            emit(tab[kSmSlots * (n - 1) + 1], true, true);
            i = 1;
            end = n - 1;
        }
    } else if (pd.flags & PD_FS) {  // the first node's split (qcgd.hpp:629-633)
        const bool take = bits & 1;
        d = 1;
        mag = cmul(mag, ct[take ? 2u : 0u]);  // a split: do_ / do_not
        if (take) {
            if (pd.flags & PD_HMLZ) {
                emit(tab[1], true, false);
                emit(tab[2], false, true);
            } else {
                overflow = true;
                emit(tab[2], false, true);
            }
        } else {
            emit(tab[0], true, true);
        }
        i = 1;
    }
    while (i < end) {
        const bool l = (lmask >> i) & 1, r = (rmask >> i) & 1;
        const bool dec = (dmask >> i) & 1;
        const bool take = dec && ((bits >> d) & 1);
        d += dec ? 1u : 0u;
        const bool split = l && r;
        if (dec) mag = cmul(mag, ct[(take ? 2u : 0u) | (split ? 0u : 1u)]);
        emit(tab[kSmSlots * i + (take ? 1u : 0u)], take || l, take ? !split : r);
        if (take && split) emit(tab[kSmSlots * i + 2], false, true);
        i += (take && !split) ? 2u : 1u;
    }
    if (overflow) {  // the left half of the first node goes last (qcgd.hpp:838-845)
        nh = step_pm(nh, tab[1]);
        lh = small_step_pm(lh, c);
    }
    return finish_pm(nh, lh, rh);
}

// (32-bit masks when every node index fits)
__device__ inline u64 sm_child(const PDesc &pd, const u64 *tab, u64 bits, const cplx *ct, cplx &mag) {
    return pd.n <= 32 ? sm_child_m<u32>(pd, tab, bits, ct, mag) : sm_child_m<u64>(pd, tab, bits, ct, mag);
}

// One child of a flip rule walked on its own, with all its decision bits
// (bit d: decision d, in node order).  flip_run's three sequences are
// independent: the magnitude takes one factor per decision node in node
// order (chosen by the node's parent left bit and the decision), and the
// left and right folds run over the nodes whose bits end up set (the
// parent's masks with the taken decision nodes flipped), in node order.
// Computed apart, each over the set bits of a mask, they are the same
// products and folds in the same order, without flip_run's per-node
// branches, which lanes holding different children would take both ways.
__device__ inline u32 ctz_mask(u32 m) { return u32(__builtin_ctz(m)); }
__device__ inline u32 ctz_mask(u64 m) { return u32(__builtin_ctzll(m)); }

// (M: the masks' word, 32 bits when every node index fits, which halves the
// bit scans' instructions)
template <typename M>
__device__ inline u64 flip_child_m(const PDesc &pd, u64 bits, const cplx *ct, cplx &mag) {
    M dm = M(pd.dmask), flips = 0;
    const M lmask = M(pd.lmask);
    u32 d = 0;
    while (dm) {
        const u32 i = ctz_mask(dm);
        dm &= dm - 1;
        const bool take = (bits >> d++) & 1;
        const u32 l = u32(lmask >> i) & 1u;
        mag = cmul(mag, ct[(take ? 2u : 0u) | l]);
        flips |= M(take) << i;
    }
    M lm = lmask ^ flips, rm = M(pd.rmask) ^ flips;
    u64 lh = 0, rh = 0;
    while (lm) {
        lh = small_step_pm(lh, u64(ctz_mask(lm)));
        lm &= lm - 1;
    }
    while (rm) {
        rh = small_step_pm(rh, u64(ctz_mask(rm)));
        rm &= rm - 1;
    }
    return finish_pm(pd.nf, lh, rh);
}

__device__ inline u64 flip_child(const PDesc &pd, u64 bits, const cplx *ct, cplx &mag) {
    return pd.n <= 32 ? flip_child_m<u32>(pd, bits, ct, mag) : flip_child_m<u64>(pd, bits, ct, mag);
}

// The children of a work item, depth first over its t = min(k, 3) trailing
// decisions s, s+1, s+2 (bit j of the child's item index u is decision
// s + j): each branch walks the nodes up to the next decision once and its
// two continuations start from there, so siblings share every node before
// the first decision in which they differ (2 + 4 + 8 segment walks in place
// of 8 whole tails).  Each child still sees the same sequence of magnitude
// products and hash folds as a walk of its own.  leaf(u, hash, magnitude)
// runs once per child, u a compile-time constant after unrolling.
static_assert(kItemBits == 3, "walk_item enumerates three decision levels");

template <int KIND, typename W, typename KC>
__device__ __attribute__((always_inline)) inline void walk_step(W &w, const PDesc &pd, const u64 *ptab, u32 bit,
                                                                u32 dlimit, const KC &k) {
    if (KIND == QD_RULE_QCGD_SPLIT_MERGE) sm_run(w, pd, ptab, bit, dlimit, k);
    else flip_run<KIND == QD_RULE_QCGD_COIN>(w, pd, bit, dlimit, k);
}

template <int KIND, typename W, typename KC, typename Leaf>
__device__ __attribute__((always_inline)) inline void walk_item(const W &pre, const PDesc &pd, const u64 *ptab, u32 s,
                                                                u32 t, const KC &k, Leaf &&leaf) {
#pragma unroll
    for (u32 b0 = 0; b0 < 2; ++b0) {
        if (b0 && t < 1) break;
        W w0 = pre;
        walk_step<KIND>(w0, pd, ptab, b0, s + 1, k);
#pragma unroll
        for (u32 b1 = 0; b1 < 2; ++b1) {
            if (b1 && t < 2) break;
            W w1 = w0;
            walk_step<KIND>(w1, pd, ptab, b1, s + 2, k);
#pragma unroll
            for (u32 b2 = 0; b2 < 2; ++b2) {
                if (b2 && t < 3) break;
                W w2 = w1;
                walk_step<KIND>(w2, pd, ptab, b2, s + 3, k);
                u64 h;
                if (KIND == QD_RULE_QCGD_SPLIT_MERGE) h = sm_finish(w2, ptab);
                else h = finish_pm(pd.nf, w2.lh, w2.rh);
                leaf(b0 | (b1 << 1) | (b2 << 2), h, w2.mag);
            }
        }
    }
}

constexpr int kGenThreads = 256;

// Parent of the first work item of every chunk of kGenThreads items (the
// chunk a generate workgroup takes per iteration): cpar[c] = the last p
// with wb[p] <= c * kGenThreads.
__global__ __launch_bounds__(256) void k_chunk_parent(const u64 *wb, u64 N, u64 nchunk, u64 *cpar) {
    for (u64 c = u64(blockIdx.x) * blockDim.x + threadIdx.x; c < nchunk; c += u64(gridDim.x) * blockDim.x)
        cpar[c] = find_parent(wb, N, c * kGenThreads);
}

// One thread per work item; writes hash and magnitude of its children at
// their candidate index (coalesced for each of the up-to-8 children) and
// histograms the top b1 hash bits for the first partition pass.
// BATCH: parents belong to instances (iseg: ninst + 1 parent offsets); each
// takes its instance's rule coefficients and adds instance * fold to its
// children's hashes.
template <int KIND, bool BATCH>
__global__ __launch_bounds__(kGenThreads) void k_generate_fast(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                              const u64 *wb, u64 N, const u64 *cpar, const u64 *tab,
                                                              const u64 *pbegin,
                                                              qd_rule rule, u64 W, u64 *H0, double2 *M0, u32 b1,
                                                              u64 *hist1, const qd_rule *rules, const u64 *iseg,
                                                              u32 ninst, u64 fold) {
    __shared__ u32 hist[kHistBins];
    __shared__ u64 swb[kGenThreads + 1];
    __shared__ u64 pfirst;
    if (b1)
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x) hist[i] = 0;
    QcgdCoef k = coef_of(rule);
    for (u64 w0 = u64(blockIdx.x) * kGenThreads; w0 < W; w0 += u64(gridDim.x) * kGenThreads) {
        __syncthreads();
        if (threadIdx.x == 0) pfirst = cpar[w0 / kGenThreads];
        __syncthreads();
        const u64 pf = pfirst;
        for (u32 q = threadIdx.x; q <= kGenThreads; q += blockDim.x) swb[q] = pf + q <= N ? wb[pf + q] : ~0ull;
        __syncthreads();
        const u64 w = w0 + threadIdx.x;
        if (w >= W) continue;
        // parent of w: the last p with wb[p] <= w
        u64 p;
        if (swb[kGenThreads] > w) {
            u32 lo = 0, hi = kGenThreads;
            while (lo + 1 < hi) {
                const u32 mid = (lo + hi) >> 1;
                if (swb[mid] <= w) lo = mid;
                else hi = mid;
            }
            p = pf + lo;
        } else {
            p = find_parent(wb, N, w);
        }
        u64 hadd = 0;
        if (BATCH) {
            u32 lo = 0, hi = ninst - 1;  // instance of parent p
            while (lo < hi) {
                const u32 mid = (lo + hi + 1) >> 1;
                if (iseg[mid] <= p) lo = mid;
                else hi = mid - 1;
            }
            k = coef_of(rules[lo]);
            hadd = u64(lo) * fold;
        }
        const PDesc pd = D[p];
        const u32 j0 = u32(w - wb[p]);
        const u32 t = min(u32(pd.k), kItemBits);
        const u32 s = pd.k - t;
        const double2 m0 = pmag[p];
        Walk pre{};
        pre.mag = cplx{m0.x, m0.y};
        const u64 *ptab = tab + sm_tab_base(pbegin[p]);
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) sm_run(pre, pd, ptab, j0, s, k);
        else flip_run<KIND == QD_RULE_QCGD_COIN>(pre, pd, j0, s, k);
        const u64 cbase = cb[p] + j0;
        walk_item<KIND>(pre, pd, ptab, s, t, k, [&](u32 u, u64 h, const cplx &m) {
            if (BATCH) h += hadd;
            const u64 c = cbase + (u64(u) << s);
            H0[c] = h;
            M0[c] = make_double2(m.re, m.im);
            if (b1) atomicAdd(&hist[h >> (64 - b1)], 1u);
        });
    }
    if (b1) {
        __syncthreads();
        for (int i = threadIdx.x; i < (1 << b1); i += blockDim.x)
            if (hist[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist1 + i), (unsigned long long)hist[i]);
    }
}

// ---------------------------------------------------------------------------
// Generate fused with the first radix pass (at most 2^9 buckets): a block
// ranks its children per bucket of the top b1 key bits in LDS, reserves one
// range per bucket and writes 32-byte entries {key, candidate, magnitude}
// into fixed-capacity buckets (bucket d owns [d * cap1, (d + 1) * cap1) of
// E1; cnt1[d] counts them).  A bucket that would pass cap1 drops the excess
// and raises ERR_SPILL1: the step then redoes generate and partition the
// two-kernel way.  BATCH as in k_generate_fast.
// ---------------------------------------------------------------------------
constexpr u32 ERR_SPILL1 = 32u;
constexpr u32 kFusedBins = 512;

template <int KIND, bool BATCH>
__global__ __launch_bounds__(kGenThreads, 3) void k_generate_scatter(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                                 const u64 *wb, u64 N, const u64 *cpar, const u64 *tab,
                                                                 const u64 *pbegin, qd_rule rule, u64 W, u32 b1,
                                                                 u64 cap1, u64 *cnt1, Entry *E1, u32 *err,
                                                                 const qd_rule *rules, const u64 *iseg, u32 ninst,
                                                                 u64 fold) {
    constexpr u32 kKids = 1u << kItemBits;
    __shared__ u32 bcnt[kFusedBins];
    __shared__ u64 bbase[kFusedBins];  // first reserved slot of the bucket
    __shared__ u64 bidx[kFusedBins];   // its entry index in E1: bucket * cap1 + first slot
    __shared__ u64 swb[kGenThreads + 1];
    __shared__ u64 pfirst;
    // batched: the children's magnitudes wait in LDS (each thread its own
    // slots) rather than registers, which the instance search and offsets
    // need
    __shared__ double2 sms[BATCH ? kKids * kGenThreads : 1];
    const u32 nbin = 1u << b1;
    QcgdCoef k = coef_of(rule);
    bool spill = false;
    for (u64 w0 = u64(blockIdx.x) * kGenThreads; w0 < W; w0 += u64(gridDim.x) * kGenThreads) {
        __syncthreads();
        if (threadIdx.x == 0) pfirst = cpar[w0 / kGenThreads];
        for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) bcnt[i] = 0;
        __syncthreads();
        const u64 pf = pfirst;
        for (u32 q = threadIdx.x; q <= kGenThreads; q += blockDim.x) swb[q] = pf + q <= N ? wb[pf + q] : ~0ull;
        __syncthreads();
        const u64 w = w0 + threadIdx.x;
        // children u < nk of this item; the arrays are indexed by the
        // unrolled loop counter only, so they stay in registers
        u64 hs[kKids];
        double2 ms[kKids];
        u32 br[kKids];
        u32 nk = 0, s = 0;
        u64 cbase = 0;
        if (w < W) {
            u64 p;
            if (swb[kGenThreads] > w) {
                u32 lo = 0, hi = kGenThreads;
                while (lo + 1 < hi) {
                    const u32 mid = (lo + hi) >> 1;
                    if (swb[mid] <= w) lo = mid;
                    else hi = mid;
                }
                p = pf + lo;
            } else {
                p = find_parent(wb, N, w);
            }
            u64 hadd = 0;
            if (BATCH) {
                u32 lo = 0, hi = ninst - 1;  // instance of parent p
                while (lo < hi) {
                    const u32 mid = (lo + hi + 1) >> 1;
                    if (iseg[mid] <= p) lo = mid;
                    else hi = mid - 1;
                }
                k = coef_of(rules[lo]);
                hadd = u64(lo) * fold;
            }
            const PDesc pd = D[p];
            const u32 j0 = u32(w - wb[p]);
            const u32 t = min(u32(pd.k), kItemBits);
            s = pd.k - t;
            nk = 1u << t;
            const double2 m0 = pmag[p];
            Walk pre{};
            pre.mag = cplx{m0.x, m0.y};
            const u64 *ptab = tab + sm_tab_base(pbegin[p]);
            if (KIND == QD_RULE_QCGD_SPLIT_MERGE) sm_run(pre, pd, ptab, j0, s, k);
            else flip_run<KIND == QD_RULE_QCGD_COIN>(pre, pd, j0, s, k);
            cbase = cb[p] + j0;
            walk_item<KIND>(pre, pd, ptab, s, t, k, [&](u32 u, u64 h, const cplx &m) {
                if (BATCH) h += hadd;
                const u32 d = b1 ? u32(h >> (64 - b1)) : 0u;
                hs[u] = h;
                if (BATCH) sms[u * kGenThreads + threadIdx.x] = make_double2(m.re, m.im);
                else ms[u] = make_double2(m.re, m.im);
                br[u] = (d << 16) | atomicAdd(&bcnt[d], 1u);
            });
        }
        __syncthreads();
        for (u32 i = threadIdx.x; i < nbin; i += blockDim.x)
            if (bcnt[i]) {
                const u64 b0 = atomicAdd(reinterpret_cast<unsigned long long *>(cnt1 + i), (unsigned long long)bcnt[i]);
                bbase[i] = b0;
                bidx[i] = u64(i) * cap1 + b0;
            }
        __syncthreads();
#pragma unroll
        for (u32 u = 0; u < kKids; ++u) {
            if (u < nk) {
                const u32 d = br[u] >> 16, r = br[u] & 0xFFFF;
                if (bbase[d] + r < cap1) {
                    const double2 mu = BATCH ? sms[u * kGenThreads + threadIdx.x] : ms[u];
                    Entry e;
                    e.hash = hs[u];
                    e.cand = cbase + (u64(u) << s);
                    e.re = mu.x;
                    e.im = mu.y;
                    E1[bidx[d] + r] = e;
                } else {
                    spill = true;
                }
            }
        }
    }
    if (spill) atomicOr(err, ERR_SPILL1);
}

// ---------------------------------------------------------------------------
// Hash-range passes, for steps whose candidates exceed the scratch memory.
// The counting pass walks every work item once: it histograms the top 12
// bits of every child hash into hist (4096 bins) and tags each child with
// its slice of the hash space, slice = bin * S >> 12 for S = 2^T - 1 equal
// slices (T <= 8); the tag S marks a child the work item does not have.
// The tags are stored as T bit planes of one byte per work item: bit u of
// plane b's byte for item w is bit b of child u's tag.  A range pass then
// lists, from the planes alone, exactly the children of the slices it
// covers and walks each of them on its own lane, so that across the passes
// every child is walked once (once per pass of its slice when a slice is
// split over passes), not once per work item that has a child in the pass.
// ---------------------------------------------------------------------------
__device__ inline u64 rotl12(u64 h) { return (h << 12) | (h >> 52); }
__device__ inline u64 rotr12(u64 k) { return (k >> 12) | (k << 52); }

constexpr u32 kMaxTagBits = 8;

template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_range_count(const PDesc *D, const double2 *pmag, const u64 *wb, u64 N,
                                                            const u64 *cpar, const u64 *tab, const u64 *pbegin,
                                                            qd_rule rule, u64 W, u32 T, u8 *planes, u64 pstride,
                                                            u64 *hist) {
    __shared__ u32 lh[kHistBins];
    __shared__ u64 swb[kGenThreads + 1];
    __shared__ u64 pfirst;
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) lh[i] = 0;
    const QcgdCoef k = coef_of(rule);
    const u32 S = (1u << T) - 1;
    for (u64 w0 = u64(blockIdx.x) * kGenThreads; w0 < W; w0 += u64(gridDim.x) * kGenThreads) {
        __syncthreads();
        if (threadIdx.x == 0) pfirst = cpar[w0 / kGenThreads];
        __syncthreads();
        const u64 pf = pfirst;
        for (u32 q = threadIdx.x; q <= kGenThreads; q += blockDim.x) swb[q] = pf + q <= N ? wb[pf + q] : ~0ull;
        __syncthreads();
        const u64 w = w0 + threadIdx.x;
        if (w >= W) continue;
        u64 p;
        if (swb[kGenThreads] > w) {
            u32 lo = 0, hi = kGenThreads;
            while (lo + 1 < hi) {
                const u32 mid = (lo + hi) >> 1;
                if (swb[mid] <= w) lo = mid;
                else hi = mid;
            }
            p = pf + lo;
        } else {
            p = find_parent(wb, N, w);
        }
        const PDesc pd = D[p];
        const u32 j0 = u32(w - wb[p]);
        const u32 t = min(u32(pd.k), kItemBits);
        const u32 s = pd.k - t;
        const double2 m0 = pmag[p];
        Walk pre{};
        pre.mag = cplx{m0.x, m0.y};
        const u64 *ptab = tab + sm_tab_base(pbegin[p]);
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) sm_run(pre, pd, ptab, j0, s, k);
        else flip_run<KIND == QD_RULE_QCGD_COIN>(pre, pd, j0, s, k);
        u32 tg[1u << kItemBits];
#pragma unroll
        for (u32 u = 0; u < (1u << kItemBits); ++u) tg[u] = S;
        walk_item<KIND>(pre, pd, ptab, s, t, k, [&](u32 u, u64 h, const cplx &) {
            const u32 bin = u32(h >> 52);
            atomicAdd(&lh[bin], 1u);
            tg[u] = (bin * S) >> 12;
        });
#pragma unroll
        for (u32 b = 0; b < kMaxTagBits; ++b) {
            if (b >= T) break;
            u32 byte = 0;
#pragma unroll
            for (u32 u = 0; u < (1u << kItemBits); ++u) byte |= ((tg[u] >> b) & 1u) << u;
            planes[u64(b) * pstride + w] = u8(byte);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
        if (lh[i]) atomicAdd(reinterpret_cast<unsigned long long *>(hist + i), (unsigned long long)lh[i]);
}

// Bit-sliced x < c over 64 tags at once: x[b] holds bit b of each tag.
__device__ inline u64 tags_below(const u64 *x, u32 T, u32 c) {
    u64 lt = 0, eq = ~0ull;
#pragma unroll
    for (int b = int(kMaxTagBits) - 1; b >= 0; --b) {
        if (u32(b) >= T) continue;
        if ((c >> b) & 1u) {
            lt |= eq & ~x[b];
            eq &= x[b];
        } else {
            eq &= ~x[b];
        }
    }
    return lt;
}

// ---------------------------------------------------------------------------
// A hash-range pass, fused with its first radix pass, in two walks without
// global atomics.  Block b of a fixed grid takes the tiles b, b + grid, ...
// of kRangeTile work items.  For each tile it lists the children whose tag
// lies in the pass's slices [slo, shi) (a thread reads the planes of 16
// consecutive items; the list is in tile order) and walks each listed child
// on its own lane.  In the first walk it counts the children whose top 12
// hash bits lie in [rlo, rhi) per bucket of the top b1 bits of their key
// (the hash rotated left by 12) and writes its counts as row b of bmat.  A
// column scan turns bmat into every block's first slot per bucket; the
// second walk visits the same children again and writes each as a 32-byte
// entry at its block's next slot of its bucket.  The buckets are exactly
// sized and lie back to back, split over up to four allocations by bucket
// (EParts: bucket d in part d >> pshift, at its offset within the part).
// ---------------------------------------------------------------------------
struct EParts {
    Entry *p[4];
    u32 pshift;  // bucket d lives in part d >> pshift
};

constexpr u32 kRangeBins = 1024;  // first-pass buckets of a range pass, at most
constexpr u32 kRangeTile = 4096;  // work items per tile
constexpr u32 kRangePer = kRangeTile / kGenThreads;  // consecutive items a thread lists (16)
constexpr u32 kRangeList = 3584;  // listed children walked per round of a tile
constexpr u32 kRangeWin = 512;    // parents of the LDS window of wb (beyond: a global search)
constexpr u32 kRangeGrid = 2048;  // blocks of the two walks (rows of bmat)
static_assert(kRangePer % 8 == 0 && kRangeTile % kGenThreads == 0, "tiles of whole chunks, 8-item plane words");
static_assert(kRangeTile * (1u << kItemBits) <= 65536, "list entries are 16-bit");

struct RangeSel {
    const u8 *planes;  // T planes of pstride bytes (pstride a multiple of kRangeTile)
    u64 pstride;
    u32 T, slo, shi;   // tag bits; the pass's slices
    u32 rlo, rhi;      // its range of the top 12 hash bits
};

struct RangeLds {
    cplx ct[4];  // the decision factors (fill_factor_table)
    u64 swb[kRangeWin + 1];
    u32 list[kRangeList];  // (parent offset in the window << 15) | (item in tile << 3) | child
    u64 tmp[kGenThreads / kWave + 1];
    u64 pfirst;
};
constexpr u32 kRangeNoWin = 0x1FFFFu;  // list entry's parent beyond the window: searched in wb

// The walks' shared structure: leaf(key, magnitude, candidate) for every
// child of the pass in the range, in a fixed order per block.
template <int KIND, class Leaf>
__device__ __attribute__((always_inline)) inline void range_walks(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                                  const u64 *wb, u64 N, const u64 *cpar,
                                                                  const u64 *tab, const u64 *pbegin,
                                                                  const QcgdCoef &k, u64 W, const RangeSel &rs,
                                                                  RangeLds &L, Leaf &&leaf) {
    if (threadIdx.x == 0) fill_factor_table(k, L.ct);  // (published by the tile loop's first barrier)
    for (u64 t0 = u64(blockIdx.x) * kRangeTile; t0 < W; t0 += u64(gridDim.x) * kRangeTile) {
        __syncthreads();
        if (threadIdx.x == 0) L.pfirst = cpar[t0 / kGenThreads];
        __syncthreads();
        const u64 pf = L.pfirst;
        for (u32 q = threadIdx.x; q <= kRangeWin; q += blockDim.x) L.swb[q] = pf + q <= N ? wb[pf + q] : ~0ull;
        // this thread's items t0 + kRangePer * tid + [0, kRangePer): 64 child
        // bits per word of 8 items (bit 8i + u: item i, child u)
        u64 in[kRangePer / 8];
        u32 cnt = 0;
#pragma unroll
        for (u32 q = 0; q < kRangePer / 8; ++q) {
            const u64 w8 = t0 + u64(threadIdx.x) * kRangePer + 8 * q;
            in[q] = 0;
            if (w8 < W) {
                u64 x[kMaxTagBits];
#pragma unroll
                for (u32 b = 0; b < kMaxTagBits; ++b)
                    x[b] = b < rs.T ? *reinterpret_cast<const u64 *>(rs.planes + u64(b) * rs.pstride + w8) : 0ull;
                in[q] = ~tags_below(x, rs.T, rs.slo) & tags_below(x, rs.T, rs.shi);
                if (W - w8 < 8) in[q] &= (1ull << (8 * (W - w8))) - 1;
            }
            cnt += u32(__popcll(in[q]));
        }
        u64 total;
        const u32 first = u32(block_excl_scan(cnt, L.tmp, total));  // (its barriers: swb is in place)
        // the parent of this thread's first item, in the window (one search;
        // its other items' parents follow by stepping forward)
        const u64 wt = t0 + u64(threadIdx.x) * kRangePer;
        u32 lp0 = kRangeNoWin;
        if (cnt && L.swb[kRangeWin] > wt + kRangePer - 1) {
            u32 lo = 0, hi = kRangeWin;
            while (lo + 1 < hi) {
                const u32 mid = (lo + hi) >> 1;
                if (L.swb[mid] <= wt) lo = mid;
                else hi = mid;
            }
            lp0 = lo;
        }
        for (u32 base = 0; base < u32(total); base += kRangeList) {
            if (base) __syncthreads();  // the previous round's list is walked
            u32 o = first, lp = lp0;
#pragma unroll
            for (u32 q = 0; q < kRangePer / 8; ++q) {
                u64 m = in[q];
                while (m) {
                    const u32 bit = u32(__builtin_ctzll(m));
                    m &= m - 1;
                    const u32 it = threadIdx.x * kRangePer + 8 * q + (bit >> 3);
                    if (lp != kRangeNoWin)
                        while (L.swb[lp + 1] <= t0 + it) ++lp;
                    if (o >= base && o < base + kRangeList) L.list[o - base] = (lp << 15) | (it << 3) | (bit & 7u);
                    ++o;
                }
            }
            __syncthreads();
            const u32 nl = min(u32(total) - base, kRangeList);
            for (u32 i = threadIdx.x; i < nl; i += kGenThreads) {
                const u32 e = L.list[i];
                const u64 w = t0 + ((e & 0x7FFFu) >> 3);
                const u32 u = e & 7u;
                const u32 lp = e >> 15;
                const u64 p = lp != kRangeNoWin ? pf + lp : find_parent(wb, N, w);
                const PDesc pd = D[p];
                const u32 t = min(u32(pd.k), kItemBits);
                if (u >> t) continue;  // tagged as absent: never listed
                const u32 s = pd.k - t;
                const u32 j0 = u32(w - wb[p]);
                const double2 m0 = pmag[p];
                Walk wk{};
                wk.mag = cplx{m0.x, m0.y};
                // the child's own branch of walk_item's tree in one walk over
                // all its decisions (the same folds): one loop over the nodes,
                // so that lanes of different parents wait for each other
                // once, not once per segment
                const u64 bits = u64(j0) | (u64(u) << s);
                u64 h;
                if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
                    h = sm_child(pd, tab + sm_tab_base(pbegin[p]), bits, L.ct, wk.mag);
                } else {
                    h = flip_child(pd, bits, L.ct, wk.mag);
                }
                const u32 top = u32(h >> 52);
                if (top >= rs.rlo && top < rs.rhi) leaf(rotl12(h), wk.mag, cb[p] + j0 + (u64(u) << s));
            }
        }
    }
}

// First walk: per-block counts of the range's children per bucket, as row
// blockIdx.x of bmat (nbin columns).
template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_range_bincount(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                                const u64 *wb, u64 N, const u64 *cpar,
                                                                const u64 *tab, const u64 *pbegin, qd_rule rule,
                                                                u64 W, RangeSel rs, u32 b1, u32 *bmat) {
    __shared__ u32 bcnt[kRangeBins];
    __shared__ RangeLds L;
    const u32 nbin = 1u << b1;
    for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) bcnt[i] = 0;
    const QcgdCoef k = coef_of(rule);
    range_walks<KIND>(D, pmag, cb, wb, N, cpar, tab, pbegin, k, W, rs, L,
                      [&](u64 key, const cplx &, u64) { atomicAdd(&bcnt[b1 ? u32(key >> (64 - b1)) : 0u], 1u); });
    __syncthreads();
    for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) bmat[u64(blockIdx.x) * nbin + i] = bcnt[i];
}

// Column scan of bmat (rows = the walks' blocks, `stride` counters apart;
// nbin columns): every block's first slot in bucket d, counted from the
// bucket's start; total[d] = the bucket's size.  One workgroup per bucket.
__global__ __launch_bounds__(256) void k_range_colscan(u32 *bmat, u32 rows, u32 nbin, u32 stride, u64 *total) {
    __shared__ u64 tmp[256 / kWave + 1];
    const u32 d = blockIdx.x;
    constexpr u32 kPer = kRangeGrid / 256;
    u32 v[kPer];
    u64 sum = 0;
#pragma unroll
    for (u32 q = 0; q < kPer; ++q) {
        const u32 r = threadIdx.x * kPer + q;
        v[q] = r < rows ? bmat[u64(r) * stride + d] : 0u;
        sum += v[q];
    }
    u64 tot;
    u64 run = block_excl_scan(sum, tmp, tot);
#pragma unroll
    for (u32 q = 0; q < kPer; ++q) {
        const u32 r = threadIdx.x * kPer + q;
        if (r < rows) bmat[u64(r) * stride + d] = u32(run);
        run += v[q];
    }
    if (threadIdx.x == 0) total[d] = tot;
}

// Second walk: the same children, each written at its block's next slot of
// its bucket (boff[d]: the bucket's first entry within its part).
template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_range_binwrite(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                                const u64 *wb, u64 N, const u64 *cpar,
                                                                const u64 *tab, const u64 *pbegin, qd_rule rule,
                                                                u64 W, RangeSel rs, u32 b1, const u32 *bmat,
                                                                const u64 *boff, EParts E) {
    __shared__ u64 cur[kRangeBins];
    __shared__ RangeLds L;
    const u32 nbin = 1u << b1;
    for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) cur[i] = boff[i] + bmat[u64(blockIdx.x) * nbin + i];
    const QcgdCoef k = coef_of(rule);
    range_walks<KIND>(D, pmag, cb, wb, N, cpar, tab, pbegin, k, W, rs, L, [&](u64 key, const cplx &m, u64 cand) {
        const u32 d = b1 ? u32(key >> (64 - b1)) : 0u;
        const u64 at = atomicAdd(reinterpret_cast<unsigned long long *>(&cur[d]), 1ull);
        Entry e;
        e.hash = key;
        e.cand = cand;
        e.re = m.re;
        e.im = m.im;
        E.p[d >> E.pshift][at] = e;
    });
}

// ---------------------------------------------------------------------------
// A hash-range pass in one walk.  Each child of the pass in the range is
// written once, into its block's region of its group (the top gbits bits of
// its key): group q's regions lie at E.g[q], block b's at b * capr, and
// rcount[q * gridDim.x + b] receives the region's count.  Places within a
// region come from wavefront-aggregated LDS counters, one per group.  A
// region that would pass capr raises ERR_SPILL1 (the caller then walks the
// pass again the two-walk way).  The block also counts its children per
// first-level bucket (the top b1 key bits) as row blockIdx.x of bmat, so
// that each group's regions merge into its first-level buckets at exact
// offsets (k_range_colscan over the group's columns, k_excl_small,
// k_region_place).
// ---------------------------------------------------------------------------
constexpr u32 kRangeGroupsMax = 8;

struct GroupBases {
    Entry *g[kRangeGroupsMax];
};

// Wavefronts per SIMD of the one-walk pass: a flip-rule walk fits 7 (72
// registers; its LDS, 22.7 KB a block, fits 7 blocks), a split_merge walk 6.
template <int KIND>
constexpr u32 kRangeWalkWaves = KIND == QD_RULE_QCGD_SPLIT_MERGE ? 6u : 7u;

template <int KIND>
__global__ __launch_bounds__(kGenThreads, kRangeWalkWaves<KIND>) void k_range_walkwrite(const PDesc *D, const double2 *pmag, const u64 *cb,
                                                                 const u64 *wb, u64 N, const u64 *cpar,
                                                                 const u64 *tab, const u64 *pbegin, qd_rule rule,
                                                                 u64 W, RangeSel rs, u32 gbits, GroupBases E,
                                                                 u64 capr, u32 *rcount, u32 b1, u32 *bmat,
                                                                 u32 *err) {
    __shared__ u32 lcnt[kRangeGroupsMax];
    __shared__ u32 bcnt[kRangeBins];
    __shared__ RangeLds L;
    const u32 nbin = 1u << b1;
    if (threadIdx.x < kRangeGroupsMax) lcnt[threadIdx.x] = 0;
    for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) bcnt[i] = 0;
    const QcgdCoef k = coef_of(rule);
    const u32 lane = threadIdx.x & (kWave - 1);
    const u32 ng = 1u << gbits;
    bool spill = false;
    range_walks<KIND>(D, pmag, cb, wb, N, cpar, tab, pbegin, k, W, rs, L, [&](u64 key, const cplx &m, u64 cand) {
        const u32 g = gbits ? u32(key >> (64 - gbits)) : 0u;
        // the active lanes of this lane's group: three ballots on its bits;
        // the group's lowest lane reserves the group's places
        u64 same = __ballot(true);
#pragma unroll
        for (u32 b = 0; b < 3; ++b) {
            if (b >= gbits) break;
            const u64 mb = __ballot((g >> b) & 1u);
            same &= ((g >> b) & 1u) ? mb : ~mb;
        }
        const u32 leader = u32(__ffsll((unsigned long long)same) - 1);
        u32 base = 0;
        if (lane == leader) base = atomicAdd(&lcnt[g], u32(__popcll(same)));
        base = __shfl(base, int(leader), kWave);
        const u32 pos = base + u32(__popcll(same & ((1ull << lane) - 1)));
        atomicAdd(&bcnt[b1 ? u32(key >> (64 - b1)) : 0u], 1u);
        if (pos < capr) {
            Entry e;
            e.hash = key;
            e.cand = cand;
            e.re = m.re;
            e.im = m.im;
            E.g[g][u64(blockIdx.x) * capr + pos] = e;
        } else {
            spill = true;
        }
    });
    __syncthreads();
    if (threadIdx.x < ng) rcount[threadIdx.x * gridDim.x + blockIdx.x] = min(lcnt[threadIdx.x], u32(capr));
    for (u32 i = threadIdx.x; i < nbin; i += blockDim.x) bmat[u64(blockIdx.x) * nbin + i] = bcnt[i];
    if (spill) atomicOr(err, ERR_SPILL1);
}

// off[i] = total[0] + ... + total[i - 1] for i <= n (n small).
__global__ void k_excl_small(const u64 *total, u32 n, u64 *off) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        u64 acc = 0;
        for (u32 i = 0; i < n; ++i) {
            off[i] = acc;
            acc += total[i];
        }
        off[n] = acc;
    }
}

// Each region's entries at its running places in the buckets: bucket d
// starts at boff[d], region r's first place in it is bm[r * bstride + d]
// after the column scan.  A chunk of kPlaceChunk entries is ordered by bucket in LDS
// first, so that consecutive lanes write consecutive places of a bucket
// (runs of about kPlaceChunk / nbin entries) rather than one entry to each
// of up to 64 buckets per store.
constexpr u32 kPlaceThreads = 1024;
constexpr u32 kPlacePer = 2;
constexpr u32 kPlaceChunk = kPlaceThreads * kPlacePer;

__global__ __launch_bounds__(kPlaceThreads) void k_region_place(const Entry *in, u64 capr, const u32 *rcount,
                                                                u32 nreg, u32 shift, u32 nbin, const u32 *bm,
                                                                u32 bstride, const u64 *boff, Entry *out) {
    // the chunk's entries by field (8-byte words: an entry's place is random,
    // and a 32-byte store per lane to random places conflicts on the LDS
    // banks far more than four 8-byte ones)
    __shared__ u64 sh[kPlaceChunk], sc[kPlaceChunk];
    __shared__ double sre[kPlaceChunk], sim[kPlaceChunk];
    __shared__ u32 cnt[kRangeBins], lst[kRangeBins];
    __shared__ u64 cur[kRangeBins];
    __shared__ u64 tmp[kPlaceThreads / kWave + 1];
    constexpr u32 kBinsPer = kRangeBins / kPlaceThreads;
    auto digit = [&](u64 key) { return nbin > 1 ? u32(key >> shift) & (nbin - 1) : 0u; };
    for (u32 r = blockIdx.x; r < nreg; r += gridDim.x) {
        __syncthreads();
        for (u32 i = threadIdx.x; i < nbin; i += kPlaceThreads) cur[i] = boff[i] + bm[u64(r) * bstride + i];
        const u32 n = rcount[r];
        const Entry *src = in + u64(r) * capr;
        for (u32 c0 = 0; c0 < n; c0 += kPlaceChunk) {
            const u32 m = min(kPlaceChunk, n - c0);
            __syncthreads();  // the previous chunk is written and its counts added
            for (u32 i = threadIdx.x; i < nbin; i += kPlaceThreads) cnt[i] = 0;
            __syncthreads();
            Entry e[kPlacePer];
            u32 dr[kPlacePer];
#pragma unroll
            for (u32 q = 0; q < kPlacePer; ++q) {
                const u32 j = threadIdx.x + kPlaceThreads * q;
                if (j < m) {
                    e[q] = src[c0 + j];
                    const u32 d = digit(e[q].hash);
                    dr[q] = (d << 16) | atomicAdd(&cnt[d], 1u);
                }
            }
            __syncthreads();
            // the buckets' starts within the chunk: an exclusive scan of cnt
            u32 v[kBinsPer];
            u64 sum = 0;
#pragma unroll
            for (u32 q = 0; q < kBinsPer; ++q) {
                const u32 b = threadIdx.x * kBinsPer + q;
                v[q] = b < nbin ? cnt[b] : 0u;
                sum += v[q];
            }
            u64 tot;
            u64 run = block_excl_scan(sum, tmp, tot);
#pragma unroll
            for (u32 q = 0; q < kBinsPer; ++q) {
                const u32 b = threadIdx.x * kBinsPer + q;
                if (b < nbin) lst[b] = u32(run);
                run += v[q];
            }
            __syncthreads();
#pragma unroll
            for (u32 q = 0; q < kPlacePer; ++q) {
                const u32 j = threadIdx.x + kPlaceThreads * q;
                if (j < m) {
                    const u32 at = lst[dr[q] >> 16] + (dr[q] & 0xFFFF);
                    sh[at] = e[q].hash;
                    sc[at] = e[q].cand;
                    sre[at] = e[q].re;
                    sim[at] = e[q].im;
                }
            }
            __syncthreads();
            for (u32 j = threadIdx.x; j < m; j += kPlaceThreads) {
                Entry x;
                x.hash = sh[j];
                x.cand = sc[j];
                x.re = sre[j];
                x.im = sim[j];
                const u32 d = digit(x.hash);
                out[cur[d] + (j - lst[d])] = x;
            }
            __syncthreads();
            for (u32 i = threadIdx.x; i < nbin; i += kPlaceThreads) cur[i] += cnt[i];
        }
    }
}

}  // namespace qd

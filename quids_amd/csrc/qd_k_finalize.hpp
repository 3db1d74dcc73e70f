#pragma once
// survivor placement, child sizes and bytes, normalisation sums, modifiers
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_rules.hpp"

namespace qd {

// ===========================================================================
// phase 7/8: order, sizes, populate
// ===========================================================================
// Survivors listed by their dedup output slot: rank = number of smaller
// representatives (popcount prefix of the representative bitmap).
// Survivors enumerated from the slot bitmap (one lane per 64-slot word, no
// bit set at or past U); each goes to the rank of its representative among
// the marked representatives.
__global__ __launch_bounds__(256) void k_place(const u64 *UR, const double2 *UM, const u64 *sbm, u64 U,
                                               const u64 *bitmap, const u64 *wprefix, u64 *orep, double2 *omag) {
    constexpr int kUnroll = 4;  // loads in flight per lane
    const u64 nw = (U + 63) / 64, stride = u64(gridDim.x) * blockDim.x;
    for (u64 w0 = u64(blockIdx.x) * blockDim.x + threadIdx.x; w0 < nw; w0 += kUnroll * stride) {
        u64 m[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u64 w = w0 + u * stride;
            m[u] = w < nw ? sbm[w] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            for (u64 mm = m[u]; mm; mm &= mm - 1) {
                const u64 i = (w0 + u * stride) * 64 + u64(__ffsll((unsigned long long)mm) - 1);
                const u64 rep = UR[i];
                const u64 word = bitmap[rep >> 6];
                const u32 bit = u32(rep & 63);
                const u64 r = wprefix[rep >> 6] + __popcll(word & ((1ull << bit) - 1));
                orep[r] = rep;
                omag[r] = UM[i];
            }
        }
    }
}

__device__ inline u32 align_pad(u32 size, u32 align) {
    if (align <= 1) return 0;
    const u32 off = align - size % align;
    return off == align ? 0 : off;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_child_size(Parents P, const u64 *cb, qd_rule rule, const u64 *orep, u64 K,
                                                    u32 align, u32 *osize, u64 *padded, u64 *opar) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 c = orep[k];
        const u64 p = find_parent(cb, P.n, c);
        opar[k] = p;  // the populate kernels start from it
        const u32 j = u32(c - cb[p]);
        const u8 *o = P.obj + P.begin[p];
        u32 s;
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
            Graph g(o);
            s = sm_child_size(g, sm_first_hmlz(g), j);
        } else {
            s = P.size[p];
        }
        osize[k] = s;
        padded[k] = u64(s) + align_pad(s, align);
    }
}

// One wavefront per survivor, everything lane-parallel:
//  * erase_create / coin / hadamard: the child is the parent with some
//    particle (qubit) bytes flipped -- one coalesced copy, the flips applied
//    in registers by the lanes whose words cover those bytes;
//  * split_merge: four survivors per wave, one 16-lane segment each: the
//    parent is staged in LDS; the segment's first lane lists the child's
//    nodes (kind, parent node) from the decisions (qcgd.hpp:850-1035); each
//    lane then sizes a child node's name tree, a segment scan places them,
//    and the lanes write the header and copy their name segments into an LDS
//    child buffer, stored with coalesced 4-byte writes.
// Objects beyond the LDS buffers (or not 4-byte aligned) are written by one
// lane straight from global memory.
constexpr int kPopParent = 1024;  // LDS bytes per split_merge survivor (four per wave)
constexpr int kPopChild = 2048;
constexpr int kPopNodes = 64;

__device__ inline u32 wave_excl_sum(u32 v, u32 &total) {
    const int lane = threadIdx.x & (kWave - 1);
    u32 inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const u32 o = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += o;
    }
    total = __shfl(inc, kWave - 1, kWave);
    return inc - v;
}

// Copies the parent's words to the child with the flips of child_id applied
// (erase_create / coin: both particle bytes of every flipped node; hadamard:
// byte `bit` xor !child_id), on a segment of SEG lanes (SEG = 64, or 32 when
// two survivors share a wavefront; the segment's lanes are active together;
// a QCGD parent has at most SEG nodes).
template <int KIND, int SEG>
__device__ inline void populate_flip_seg(const u8 *o, u32 psz, u8 *d, u64 span, u32 j, const qd_rule &rule) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u32 sl = lane & (SEG - 1), sbase = lane & ~(SEG - 1);
    const u64 smask = SEG == kWave ? ~0ull : (((1ull << SEG) - 1) << sbase);
    u64 flipmask = 0;  // nodes whose particles flip (n <= SEG)
    u32 n = 0;
    if (KIND != QD_RULE_HADAMARD) {
        n = ld16(o);
        const bool l = sl < n && o[2 + sl], r = sl < n && o[2 + n + sl];
        const bool elig = sl < n && flip_eligible(KIND == QD_RULE_QCGD_COIN, l, r);
        const u64 em = (u64(__ballot(elig)) & smask) >> sbase;
        const u32 rank = __popcll(em & ((1ull << sl) - 1));
        flipmask = (u64(__ballot(elig && rank < 32 && ((j >> rank) & 1))) & smask) >> sbase;
    }
    for (u64 w = sl; w < (span + 3) / 4; w += SEG) {
        const u32 at = u32(4 * w);
        u32 v = 0;
        if (at + 4 <= psz) v = *reinterpret_cast<const u32 *>(o + at);
        else
            for (u32 q = 0; q < 4; ++q)
                if (at + q < psz) v |= u32(o[at + q]) << (8 * q);
        for (u32 q = 0; q < 4; ++q) {
            const u32 pos = at + q;
            bool flip = false;
            if (KIND == QD_RULE_HADAMARD) flip = pos == rule.bit && j == 0;
            else if (pos >= 2 && pos < 2 + 2 * n) {
                const u32 node = (pos - 2) < n ? pos - 2 : pos - 2 - n;
                flip = (flipmask >> node) & 1;
            }
            if (flip) {
                const u32 byte = (v >> (8 * q)) & 0xFF;
                const u32 nb = KIND == QD_RULE_HADAMARD ? (byte ^ 1u) : (byte ? 0u : 1u);
                v = (v & ~(0xFFu << (8 * q))) | (nb << (8 * q));
            }
        }
        if (at + 4 <= span) *reinterpret_cast<u32 *>(d + at) = v;
        else
            for (u32 q = 0; q < 4 && at + q < span; ++q) d[at + q] = u8(v >> (8 * q));
    }
}

// One survivor, with the lanes of its segment (SEG as above); objects that
// are not 4-byte aligned or have more nodes than the segment are written by
// the segment's first lane.
template <int KIND, int SEG>
__device__ inline void populate_flip_one(const Parents &P, const u64 *cb, const qd_rule &rule, const u64 *orep,
                                         const u64 *opar, u64 k, const u64 *obegin, u8 *oobj) {
    const u32 sl = threadIdx.x & (SEG - 1);
    const u64 c = orep[k];
    const u64 p = opar[k];
    const u32 j = u32(c - cb[p]);
    const u8 *o = P.obj + P.begin[p];
    const u32 psz = P.size[p];
    const u64 ob = obegin[k], oe = obegin[k + 1];
    const u64 span = oe - ob;  // child bytes + zero padding
    u8 *d = oobj + ob;
    const bool aligned = ((P.begin[p] | ob) & 3) == 0;
    const u32 n = KIND == QD_RULE_HADAMARD ? 0 : ld16(o);
    if (aligned && n <= u32(SEG)) {
        populate_flip_seg<KIND, SEG>(o, psz, d, span, j, rule);
    } else if (sl == 0) {
        if (KIND == QD_RULE_HADAMARD) hadamard_child_write(o, psz, rule.bit, j, d);
        else flip_child_write(o, psz, KIND == QD_RULE_QCGD_COIN, j, d);
        for (u64 q = psz; q < span; ++q) d[q] = 0;
    }
}

// erase_create / coin / hadamard: two survivors per wavefront in 32-lane
// segments when both parents have at most 32 nodes, else one after the
// other on the whole wavefront; four wavefronts per block, no LDS.
template <int KIND>
__global__ __launch_bounds__(256) void k_populate_flip(Parents P, const u64 *cb, qd_rule rule, const u64 *orep,
                                                       const u64 *opar, u64 K, const u64 *obegin, const u32 *osize,
                                                       u8 *oobj) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u64 wave = (u64(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const u64 nwaves = u64(gridDim.x) * blockDim.x / kWave;
    for (u64 k0 = 2 * wave; k0 < K; k0 += 2 * nwaves) {
        u32 nq = 0;  // node count of survivor k0 + lane's parent (lanes 0..1)
        if (KIND != QD_RULE_HADAMARD && lane < 2 && k0 + lane < K) nq = ld16(P.obj + P.begin[opar[k0 + lane]]);
        if (__ballot(nq > 32) == 0) {
            const u64 k = k0 + lane / 32;
            if (k < K) populate_flip_one<KIND, 32>(P, cb, rule, orep, opar, k, obegin, oobj);
        } else {
            for (u64 k = k0; k < min(k0 + 2, K); ++k) populate_flip_one<KIND, kWave>(P, cb, rule, orep, opar, k, obegin, oobj);
        }
    }
    (void)osize;
}

// split_merge: kPopSegs survivors per single-wave block, one segment of
// kPopSeg lanes each (see above).
constexpr int kPopSeg = 16;
constexpr int kPopSegs = kWave / kPopSeg;

// Exclusive sum over the kPopSeg lanes of this lane's segment.
__device__ inline u32 seg_excl_sum(u32 v, u32 &total) {
    const int sl = threadIdx.x & (kPopSeg - 1);
    u32 inc = v;
#pragma unroll
    for (int d = 1; d < kPopSeg; d <<= 1) {
        const u32 o = __shfl_up(inc, d, kPopSeg);
        if (sl >= d) inc += o;
    }
    total = __shfl(inc, kPopSeg - 1, kPopSeg);
    return inc - v;
}

__global__ __launch_bounds__(kWave) void k_populate_sm(Parents P, const u64 *cb, qd_rule rule, const u64 *orep,
                                                    const u64 *opar, u64 K, const u64 *obegin, const u32 *osize,
                                                    u8 *oobj) {
    __shared__ __align__(16) u8 pbuf[kPopSegs][kPopParent];
    __shared__ __align__(16) u8 cbuf[kPopSegs][kPopChild];
    __shared__ u8 ckind[kPopSegs][kPopNodes], cl[kPopSegs][kPopNodes], cr[kPopSegs][kPopNodes];
    __shared__ u16 cnode[kPopSegs][kPopNodes];
    __shared__ u32 nnodes[kPopSegs];
    const u32 sg = threadIdx.x / kPopSeg, sl = threadIdx.x & (kPopSeg - 1);
    u8 *const pb = pbuf[sg];
    u8 *const cbf = cbuf[sg];
    for (u64 k0 = u64(blockIdx.x) * kPopSegs; k0 < K; k0 += u64(gridDim.x) * kPopSegs) {
        // every lane of a segment holds its survivor's scalars
        const u64 k = k0 + sg;
        const bool valid = k < K;
        u32 j = 0, psz = 0, csz = 0;
        u64 ob = 0, span = 0;
        const u8 *o = P.obj;
        bool fast = false;
        if (valid) {
            const u64 c = orep[k];
            const u64 p = opar[k];
            j = u32(c - cb[p]);
            o = P.obj + P.begin[p];
            psz = P.size[p];
            csz = osize[k];
            ob = obegin[k];
            span = obegin[k + 1] - ob;
            fast = ((P.begin[p] | ob) & 3) == 0 && psz <= u32(kPopParent) && csz <= u32(kPopChild);
        }
        u8 *d = oobj + ob;
        if (fast)
            for (u32 w = sl; w < psz / 4; w += kPopSeg)
                reinterpret_cast<u32 *>(pb)[w] = *reinterpret_cast<const u32 *>(o + 4 * w);
        __syncthreads();
        // one lane per segment lists the child's nodes
        if (fast && sl == 0) {
            Graph g(pb);
            u32 cnt = 0;
            cplx dm{0, 0};
            QcgdCoef kk{};
            sm_walk(g, sm_first_hmlz(g), j, kk, dm, false, [&](int kind, u32 i, bool l, bool r) {
                if (cnt < kPopNodes) {
                    ckind[sg][cnt] = u8(kind);
                    cnode[sg][cnt] = u16(i);
                    cl[sg][cnt] = l;
                    cr[sg][cnt] = r;
                }
                ++cnt;
            });
            nnodes[sg] = cnt;
        }
        __syncthreads();
        const u32 nc = fast ? nnodes[sg] : 0u;
        if (nc > u32(kPopNodes)) fast = false;  // more child nodes than the list holds
        if (valid && !fast && sl == 0) {
            Graph g(o);
            const u32 sz = sm_child_write(g, sm_first_hmlz(g), j, d);
            for (u64 q = sz; q < span; ++q) d[q] = 0;
        }
        if (fast) {
            // name tree of every child node: [src_b, src_e) copied after an
            // optional new root
            Graph g(pb);
            u32 base = 0;
            for (u32 c0 = 0; c0 < nc; c0 += kPopSeg) {
                const u32 cn = c0 + sl;
                u32 len = 0, sb = 0, se = 0, sb2 = 0, se2 = 0;
                bool newroot = false;
                i16 nh = 0, nt = 0;
                u64 nhash = 0;
                if (cn < nc) {
                    const u32 kind = ckind[sg][cn], i = cnode[sg][cn];
                    const u32 b = g.name_begin(i), e = g.name_begin(i + 1), m = e - b;
                    if (kind == 0) {
                        sb = b;
                        se = e;
                    } else if (kind == 1 || kind == 2) {
                        const i16 t = g.sub_type(b);
                        if (t >= 0) {
                            sb = kind == 1 ? b + 1 : b + u32(t);
                            se = kind == 1 ? b + u32(t) : e;
                        } else {
                            newroot = true;
                            nt = kind == 1 ? QCGD_DOT_L : QCGD_DOT_R;
                            nh = (nt == QCGD_DOT_L && g.sub_hmlz(b) < 0) ? i16(-1) : i16(1);
                            nhash = kind == 1 ? split_left_hash(g, b) : split_right_hash(g, b);
                            sb = b;
                            se = e;
                        }
                    } else {
                        const u32 jn = i + 1 < g.n ? i + 1 : 0;
                        const u32 bb = g.name_begin(jn), eb = g.name_begin(jn + 1);
                        if (merge_collapses(g, b, bb)) {
                            sb = b + 1;
                            se = e;
                        } else {
                            newroot = true;
                            nt = i16(m + 1);
                            nh = (g.sub_hmlz(b) < 0 || g.sub_hmlz(bb) < 0) ? i16(-1) : i16(1);
                            nhash = hash_combine(g.sub_hash(b), g.sub_hash(bb));
                            sb = b;
                            se = e;
                            sb2 = bb;
                            se2 = eb;
                        }
                    }
                    len = u32(newroot) + (se - sb) + (se2 - sb2);
                }
                u32 tot;
                const u32 off = base + seg_excl_sum(len, tot);
                if (cn < nc) {
                    u8 *hdr = cbf;
                    hdr[2 + cn] = cl[sg][cn];
                    hdr[2 + nc + cn] = cr[sg][cn];
                    st16(hdr + 2 + 2 * nc + 2 * cn, u16(off));
                    u8 *dst = cbf + 4 + 4 * nc + 16 * size_t(off);
                    if (newroot) {
                        put_sub(dst, nh, nt, nhash);
                        dst += 16;
                    }
                    const u32 *src = reinterpret_cast<const u32 *>(g.sub(sb));
                    for (u32 q = 0; q < 4 * (se - sb); ++q) reinterpret_cast<u32 *>(dst)[q] = src[q];
                    dst += 16 * (se - sb);
                    if (se2 > sb2) {
                        const u32 *src2 = reinterpret_cast<const u32 *>(g.sub(sb2));
                        for (u32 q = 0; q < 4 * (se2 - sb2); ++q) reinterpret_cast<u32 *>(dst)[q] = src2[q];
                    }
                }
                base += tot;
            }
            if (sl == 0) {
                st16(cbf, u16(nc));
                st16(cbf + 2 + 2 * nc + 2 * nc, u16(base));
            }
        }
        __syncthreads();
        if (fast) {
            for (u64 w = sl; w < (span + 3) / 4; w += kPopSeg) {
                const u64 at = 4 * w;
                if (at + 4 <= span) {
                    u32 v = 0;
                    if (at + 4 <= csz) v = reinterpret_cast<const u32 *>(cbf)[w];
                    else
                        for (u32 q = 0; q < 4; ++q)
                            if (at + q < csz) v |= u32(cbf[at + q]) << (8 * q);
                    *reinterpret_cast<u32 *>(d + at) = v;
                } else {
                    for (u64 q = at; q < span; ++q) d[q] = q < csz ? cbf[q] : 0;
                }
            }
        }
        __syncthreads();  // the buffers are reused by the next survivors
    }
}

// ===========================================================================
// phase 9: canonical sum of |m|^2 and scaling
// ===========================================================================
constexpr int kSumThreads = 256;
constexpr int kSumChunk = 2048;

// One block per chunk of 2048 values: lane t sums t, t+256, ..., t+1792 in
// order, then a halving tree s[t] += s[t+stride], stride = 128 .. 1.
__global__ __launch_bounds__(kSumThreads) void k_sum_chunks(const double2 *mag, const double *vals, u64 n,
                                                            double *out) {
    __shared__ double s[kSumThreads];
    const u64 base = u64(blockIdx.x) * kSumChunk;
    double a = 0.0;
    for (int j = 0; j < kSumChunk / kSumThreads; ++j) {
        const u64 i = base + threadIdx.x + u64(kSumThreads) * j;
        if (i < n) {
            double v;
            if (mag) {
                const double2 m = mag[i];
                v = m.x * m.x + m.y * m.y;
            } else {
                v = vals[i];
            }
            a += v;
        }
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_scale(double2 *mag, u64 n, double f) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        double2 m = mag[i];
        m.x = m.x / f;
        m.y = m.y / f;
        mag[i] = m;
    }
}

// ===========================================================================
// modifiers
// ===========================================================================
__global__ __launch_bounds__(256) void k_modifier(u8 *obj, const u64 *begin, const u32 *size, double2 *mag, u64 n,
                                                  qd_modifier m, u32 *err) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) {
        u8 *o = obj + begin[i];
        const u32 sz = size[i];
        if (m.kind <= QD_MOD_ZGATE) {
            if (m.bit >= sz || (m.kind == QD_MOD_CNOT && m.control_bit >= sz)) {
                atomicOr(err, ERR_BIT);
                continue;
            }
        } else if (!qcgd_valid(o, sz)) {
            atomicOr(err, ERR_OBJECT);
            continue;
        }
        double2 mm = mag[i];
        cplx c{mm.x, mm.y};
        modifier_apply(m, o, sz, c);
        mag[i] = make_double2(c.re, c.im);
    }
}


// Plain streaming copy, 16 bytes per thread per iteration.
__global__ __launch_bounds__(256) void k_copy_probe(const uint4 *src, uint4 *dst, u64 n) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) dst[i] = src[i];
}

}  // namespace qd

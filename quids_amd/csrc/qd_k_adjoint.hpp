#pragma once
// Adjoint (vector-Jacobian product) of one simulate() step.
//
// Forward, per survivor k: y_k = sum over the candidates (p, j) whose child
// hash is hash_k of x_p * f_pj, where f_pj is the product of the rule's
// factors for child j of parent p; then y' = y / sqrt(sum |y|^2).
// Given the cotangent g' of y' (g = dL/dRe + i dL/dIm):
//   g_y   = (g' - alpha y') / s,  alpha = Re sum conj(g'_k) y'_k,  s^2 = sum |y|^2
//   g_x_p = sum_j conj(f_pj) g_y[k(p, j)]
//   dL/dt = sum_pj Re(conj(g_y[k(p, j)]) x_p df_pj/dt)
// k(p, j) is found by looking the child's hash up in a table of the output
// state's hashes: the transpose of the dedup's scatter-add is a gather.
// Children that were cut (tolerance or top-K) are absent and contribute
// nothing: the truncation is held as a fixed mask.
#include <hip/hip_runtime.h>

#include "qd_device.hpp"
#include "qd_k_fastgen.hpp"
#include "qd_k_generate.hpp"
#include "qd_rules.hpp"

namespace qd {

constexpr u64 kEmptyKey = ~0ull;

__device__ inline u64 adj_slot(u64 h, u64 mask) { return (h ^ (h >> 31) ^ (h >> 47)) & mask; }

// Table of the output state's object hashes -> index.  Hashes are unique
// in a state (the dedup merged equal ones).  A hash equal to the empty key
// is kept aside in *special (index + 1).
template <int KIND>
__global__ __launch_bounds__(256) void k_adj_table(const u8 *obj, const u64 *begin, const u32 *size, u64 K,
                                                   qd_rule rule, u64 *keys, u32 *vals, u64 mask, u64 *special) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x) {
        const u64 h = rule_hash_object(rule, obj + begin[k], size[k]);
        if (h == kEmptyKey) {
            special[0] = k + 1;
            continue;
        }
        u64 s = adj_slot(h, mask);
        for (;;) {
            const u64 prev = atomicCAS(reinterpret_cast<unsigned long long *>(keys + s), kEmptyKey, h);
            if (prev == kEmptyKey || prev == h) {
                vals[s] = u32(k);
                break;
            }
            s = (s + 1) & mask;
        }
    }
}

__device__ inline bool adj_lookup(const u64 *keys, const u32 *vals, u64 mask, u64 special, u64 h, u32 &k) {
    if (h == kEmptyKey) {
        k = u32(special - 1);
        return special != 0;
    }
    u64 s = adj_slot(h, mask);
    for (;;) {
        const u64 key = keys[s];
        if (key == h) {
            k = vals[s];
            return true;
        }
        if (key == kEmptyKey) return false;
        s = (s + 1) & mask;
    }
}

// alpha terms Re(conj(g'_k) y'_k)
__global__ __launch_bounds__(256) void k_adj_dot(const double2 *gp, const double2 *yp, u64 K, double *out) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x)
        out[k] = gp[k].x * yp[k].x + gp[k].y * yp[k].y;
}

// g_y = (g' - alpha y') / s
__global__ __launch_bounds__(256) void k_adj_gy(const double2 *gp, const double2 *yp, u64 K, double alpha, double s,
                                                double2 *gy) {
    for (u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x; k < K; k += u64(gridDim.x) * blockDim.x)
        gy[k] = make_double2((gp[k].x - alpha * yp[k].x) / s, (gp[k].y - alpha * yp[k].y) / s);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// One wavefront per parent; lane l takes children l, l+64, ...  Per-parent
// results are reduced over the wave in a fixed butterfly order, so the
// outputs do not depend on scheduling.
template <int KIND>
__global__ __launch_bounds__(256) void k_adjoint(Parents P, const u64 *cb, qd_rule rule, DualCoef dk, int ndir,
                                                 const u64 *pre, const u64 *tab, const u64 *keys, const u32 *vals,
                                                 u64 mask, u64 special, const double2 *gy, double2 *gx,
                                                 double *gpart) {
    const u32 lane = threadIdx.x & (kWave - 1);
    const u64 wave = (u64(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const u64 nwaves = u64(gridDim.x) * blockDim.x / kWave;
    for (u64 p = wave; p < P.n; p += nwaves) {
        const u64 c0 = cb[p];
        const u32 nc = u32(cb[p + 1] - c0);
        const double2 xm = P.mag[p];
        const cplx x{xm.x, xm.y};
        const u8 *o = P.obj + P.begin[p];
        double axr = 0, axi = 0, at[kMaxDir] = {};
        for (u32 j = lane; j < nc; j += kWave) {
            u64 h;
            DualC f = dual_one();
            if (KIND == QD_RULE_HADAMARD) {
                cplx m{1.0, 0.0};
                h = hadamard_child_hash(o, P.size[p], rule.bit, rule.coef, j, m);
                f.v = m;
            } else if (KIND == QD_RULE_QCGD_SPLIT_MERGE) {
                Graph g(o);
                h = sm_child_hash(g, pre[p] != 0, tab + sm_tab_base(P.begin[p]), j, dk, f);
            } else {
                Graph g(o);
                h = flip_child_hash(g, KIND == QD_RULE_QCGD_COIN, pre[p], j, dk, f);
            }
            u32 k;
            if (!adj_lookup(keys, vals, mask, special, h, k)) continue;
            const double2 gk = gy[k];
            // conj(f) g
            axr += f.v.re * gk.x + f.v.im * gk.y;
            axi += f.v.re * gk.y - f.v.im * gk.x;
            for (int t = 0; t < kMaxDir; ++t) {
                if (t >= ndir) break;
                const cplx d = cmul(x, f.t[t]);
                at[t] += gk.x * d.re + gk.y * d.im;  // Re(conj(g) d)
            }
        }
        axr = wave_sum(axr);
        axi = wave_sum(axi);
        for (int t = 0; t < kMaxDir; ++t) at[t] = wave_sum(at[t]);
        if (lane == 0) {
            gx[p] = make_double2(axr, axi);
            for (int t = 0; t < ndir; ++t) gpart[u64(t) * P.n + p] = at[t];
        }
    }
}

// The same gather over the QCGD fast walk: one thread per work item (up to
// 8 sibling children sharing a walk prefix, enumerated depth first as in
// the generate kernels), the children's rule factors f as dual numbers
// starting from 1 (multiplied in the same order as the plain magnitudes),
// the parent magnitude x applied at the end.  A work item writes its sums
// conj(f) g and Re(conj(g) x df/dt) to wpart[w * kAdjPart ..]; then
// k_adjoint_reduce folds each parent's work items in order, so the outputs
// do not depend on scheduling.
constexpr int kAdjPart = 2 + kMaxDir;

template <int KIND>
__global__ __launch_bounds__(kGenThreads) void k_adjoint_fast(const PDesc *D, const double2 *pmag, const u64 *wb, u64 N,
                                                             const u64 *cpar, const u64 *tab, const u64 *pbegin,
                                                             DualCoef dk, int ndir, const u64 *keys, const u32 *vals,
                                                             u64 mask, u64 special, const double2 *gy, u64 W,
                                                             double *wpart) {
    __shared__ u64 swb[kGenThreads + 1];
    __shared__ u64 pfirst;
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
        const double2 xm = pmag[p];
        const cplx x{xm.x, xm.y};
        WalkT<DualC> pre{};
        pre.mag = dual_one();
        const u64 *ptab = tab + sm_tab_base(pbegin[p]);
        if (KIND == QD_RULE_QCGD_SPLIT_MERGE) sm_run(pre, pd, ptab, j0, s, dk);
        else flip_run<KIND == QD_RULE_QCGD_COIN>(pre, pd, j0, s, dk);
        double axr = 0, axi = 0, at[kMaxDir] = {};
        walk_item<KIND>(pre, pd, ptab, s, t, dk, [&](u32, u64 h, const DualC &f) {
            u32 k;
            if (!adj_lookup(keys, vals, mask, special, h, k)) return;
            const double2 gk = gy[k];
            // conj(f) g
            axr += f.v.re * gk.x + f.v.im * gk.y;
            axi += f.v.re * gk.y - f.v.im * gk.x;
#pragma unroll
            for (int d = 0; d < kMaxDir; ++d) {
                if (d < ndir) {
                    const cplx xd = cmul(x, f.t[d]);
                    at[d] += gk.x * xd.re + gk.y * xd.im;  // Re(conj(g) x df)
                }
            }
        });
        double *o = wpart + w * kAdjPart;
        o[0] = axr;
        o[1] = axi;
#pragma unroll
        for (int d = 0; d < kMaxDir; ++d) o[2 + d] = at[d];
    }
}

// Per parent: its work items' partial sums in order.
__global__ __launch_bounds__(256) void k_adjoint_reduce(const u64 *wb, u64 N, const double *wpart, int ndir,
                                                        double2 *gx, double *gpart) {
    for (u64 p = u64(blockIdx.x) * blockDim.x + threadIdx.x; p < N; p += u64(gridDim.x) * blockDim.x) {
        double a[kAdjPart] = {};
        for (u64 w = wb[p]; w < wb[p + 1]; ++w) {
#pragma unroll
            for (int i = 0; i < kAdjPart; ++i) a[i] += wpart[w * kAdjPart + i];
        }
        gx[p] = make_double2(a[0], a[1]);
#pragma unroll
        for (int d = 0; d < kMaxDir; ++d)
            if (d < ndir) gpart[u64(d) * N + p] = a[2 + d];
    }
}

// Cotangent through a modifier: g_before = conj(factor) g_after, obj being
// the state after the modifier.
__global__ __launch_bounds__(256) void k_modifier_adjoint(const u8 *obj, const u64 *begin, const u32 *size, u64 N,
                                                          qd_modifier m, double2 *g, u32 *err) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += u64(gridDim.x) * blockDim.x) {
        if ((m.kind == QD_MOD_YGATE || m.kind == QD_MOD_ZGATE) && m.bit >= size[i]) {
            atomicOr(err, ERR_BIT);
            continue;
        }
        const cplx f = modifier_factor_after(m, obj + begin[i]);
        const double2 v = g[i];
        g[i] = make_double2(f.re * v.x + f.im * v.y, f.re * v.y - f.im * v.x);
    }
}

// Observables of average_value (src/quids.hpp:208-234) for QCGD objects:
// 0 = number of nodes, 1 = particle count sum(L + R), 2 = [nodes == arg].
__device__ inline double qcgd_observable(const u8 *o, u32 size, int obs, int arg) {
    u32 n = size >= 4 ? ld16(o) : 0;
    if (4 + 4 * u64(n) > size) n = 0;  // not a QCGD object: observable 0
    if (obs == 0) return double(n);
    if (obs == 2) return n == u32(arg) ? 1.0 : 0.0;
    u32 c = 0;
    for (u32 i = 0; i < 2 * n; ++i) c += o[2 + i] != 0;
    return double(c);
}

// Per object: obs * |m|^2 into vals, and the cotangent 2 obs m of the loss
// sum obs |m|^2 with respect to m (when cot is not null).
__global__ __launch_bounds__(256) void k_observable(const u8 *obj, const u64 *begin, const u32 *size,
                                                    const double2 *mag, u64 N, int obs, int arg, double *vals,
                                                    double2 *cot) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += u64(gridDim.x) * blockDim.x) {
        const double v = qcgd_observable(obj + begin[i], size[i], obs, arg);
        const double2 m = mag[i];
        vals[i] = v * (m.x * m.x + m.y * m.y);
        if (cot) cot[i] = make_double2(2.0 * v * m.x, 2.0 * v * m.y);
    }
}

}  // namespace qd

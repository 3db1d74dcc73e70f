// C entry points of the test rules (user_rules.hpp) for the Python tests:
// each fills a qd_user_rule table whose callbacks launch the rule's kernels.
#include <deque>

#include "user_rules.hpp"

namespace {
std::deque<user_hadamard> g_h;  // the rules' storage: self of their tables
std::deque<user_ry> g_ry;
}  // namespace

extern "C" {

// kind 0: user_hadamard(bit, a); kind 1: user_ry(bit, a, b)
int qdt_user_rule(int kind, uint32_t bit, double a, double b, qd_user_rule *out) {
    if (!out) return QD_ERR_INVALID;
    if (kind == 0) {
        g_h.push_back(user_hadamard{bit, a});
        *out = quids::make_user_rule(&g_h.back());
    } else if (kind == 1) {
        g_ry.push_back(user_ry{bit, a, b});
        *out = quids::make_user_rule(&g_ry.back());
    } else {
        return QD_ERR_INVALID;
    }
    return QD_OK;
}

// obs (device, one double per object) = observable `kind` of every object
// of s: 0 = QCGD node count, 1 = set qubit bytes
int qdt_observable(int kind, qd_ctx *ctx, qd_state *s, double *obs) {
    if (kind == 0) return quids::eval_observable(ctx, s, obs_num_nodes{}, obs);
    if (kind == 1) return quids::eval_observable(ctx, s, obs_ones{}, obs);
    return QD_ERR_INVALID;
}

}  // extern "C"

// quantum_computer rules of the drop-in C++ API (reference:
// src/rules/quantum_computer.hpp): the hadamard rule and the cnot / X / Y / Z
// modifiers, as device functors of libquids_hip.
#pragma once

#include <cmath>
#include <iostream>

#include "../quids.hpp"

namespace quids::rules::quantum_computer {

namespace utils {
/// quantum_computer.hpp:10-22
inline void print(quids::it_t const &iter) {
    for (size_t oid = 0; oid < iter.num_object; ++oid) {
        uint size;
        mag_t mag;
        char const *begin;
        iter.get_object(oid, begin, size, mag);
        std::cout << "\t" << mag.real() << (mag.imag() < 0 ? " - " : " + ") << std::abs(mag.imag()) << "i  ";
        for (auto it = begin; it != begin + size; ++it) std::cout << (*it ? '1' : '0');
        std::cout << "\n";
    }
}
}  // namespace utils

/// quantum_computer.hpp:25-29
inline modifier_t cnot(uint32_t control_bit, uint32_t bit) {
    qd_modifier d{QD_MOD_CNOT, control_bit, bit, 0};
    return modifier_t(d);
}

/// quantum_computer.hpp:31-50
class hadamard : public device_rule_base {
public:
    hadamard(size_t bit_) {
        desc_.kind = QD_RULE_HADAMARD;
        desc_.bit = uint32_t(bit_);
        desc_.coef = 1 / std::sqrt(2.0);
    }
};

/// quantum_computer.hpp:52-56
inline modifier_t Xgate(size_t bit) { return modifier_t(qd_modifier{QD_MOD_XGATE, 0, uint32_t(bit), 0}); }
/// quantum_computer.hpp:58-66
inline modifier_t Ygate(size_t bit) { return modifier_t(qd_modifier{QD_MOD_YGATE, 0, uint32_t(bit), 0}); }
/// quantum_computer.hpp:68-75 (flips the bit too, as the reference does)
inline modifier_t Zgate(size_t bit) { return modifier_t(qd_modifier{QD_MOD_ZGATE, 0, uint32_t(bit), 0}); }

}  // namespace quids::rules::quantum_computer

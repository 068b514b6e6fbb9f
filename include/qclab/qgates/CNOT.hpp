// CNOT: another name of the controlled X gate (reference include/qclab/qgates/CNOT.hpp).
#pragma once

#include "qclab/qgates/CX.hpp"

namespace qclab {
namespace qgates {

template <typename T>
using CNOT = CX<T>;

}  // namespace qgates
}  // namespace qclab

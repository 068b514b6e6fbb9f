// Base class of 2-qubit rotation gates
// (reference include/qclab/qgates/QRotationGate2.hpp).
#pragma once

#include <algorithm>

#include "qclab/QAdjustable.hpp"
#include "qclab/QRotation.hpp"
#include "qclab/qgates/QGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class QRotationGate2 : public QGate2<T>, public QAdjustable {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  QRotationGate2() : QAdjustable(false), qubits_{0, 1}, rotation_() {}
  QRotationGate2(const int* qubits, const rotation_type& rot, const bool fixed = false)
      : QAdjustable(fixed), rotation_(rot) { setQubits(qubits); }
  QRotationGate2(const int* qubits, const real_type theta, const bool fixed = false)
      : QAdjustable(fixed), rotation_(theta) { setQubits(qubits); }
  QRotationGate2(const int* qubits, const real_type cos, const real_type sin, const bool fixed = false)
      : QAdjustable(fixed), rotation_(cos, sin) { setQubits(qubits); }
  QRotationGate2(const int qubit0, const int qubit1, const rotation_type& rot, const bool fixed = false)
      : QAdjustable(fixed), rotation_(rot) { const int q[2] = {qubit0, qubit1}; setQubits(q); }
  QRotationGate2(const int qubit0, const int qubit1, const real_type theta, const bool fixed = false)
      : QAdjustable(fixed), rotation_(theta) { const int q[2] = {qubit0, qubit1}; setQubits(q); }
  QRotationGate2(const int qubit0, const int qubit1, const real_type cos, const real_type sin,
                 const bool fixed = false)
      : QAdjustable(fixed), rotation_(cos, sin) { const int q[2] = {qubit0, qubit1}; setQubits(q); }

  bool fixed() const override { return QAdjustable::fixed(); }
  bool controlled() const override { return false; }
  int qubit() const override { return qubits_[0]; }
  std::vector<int> qubits() const override { return {qubits_[0], qubits_[1]}; }
  void setQubits(const int* qubits) override {
    assert(qubits[0] >= 0 && qubits[1] >= 0 && qubits[0] != qubits[1]);
    qubits_[0] = std::min(qubits[0], qubits[1]);
    qubits_[1] = std::max(qubits[0], qubits[1]);
  }

  const rotation_type& rotation() const { return rotation_; }
  real_type theta() const { return rotation_.theta(); }
  real_type cos() const { return rotation_.cos(); }
  real_type sin() const { return rotation_.sin(); }
  void update(const rotation_type& rot) { assert(!fixed()); rotation_ = rot; }
  void update(const real_type theta) { assert(!fixed()); rotation_.update(theta); }
  void update(const real_type cos, const real_type sin) { assert(!fixed()); rotation_.update(cos, sin); }

 protected:
  qsv_gate rotation_record(int kind, Op op, int offset) const {
    qsv_gate g = backend::make(kind, qubits_[0] + offset, qubits_[1] + offset, char(op));
    g.theta = double(theta());
    backend::set_cs(g, double(cos()), double(sin()));
    return g;
  }
  int qubits_[2];
  rotation_type rotation_;
};

}  // namespace qgates
}  // namespace qclab

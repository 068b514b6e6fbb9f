// QAdjustable: fixed/variable flag of parametrised objects
// (reference include/qclab/QAdjustable.hpp).
#pragma once

namespace qclab {

class QAdjustable {
 public:
  QAdjustable(const bool fixed = false) : fixed_(fixed) {}
  bool fixed() const { return fixed_; }
  bool variable() const { return !fixed_; }
  void makeFixed() { fixed_ = true; }
  void makeVariable() { fixed_ = false; }

 private:
  bool fixed_;
};

}  // namespace qclab

// Every public header of include/qclab, compiled by g++ (syntax only).
#include "qclab/QAdjustable.hpp"
#include "qclab/QAngle.hpp"
#include "qclab/QCircuit.hpp"
#include "qclab/QObject.hpp"
#include "qclab/QRotation.hpp"
#include "qclab/dense/SquareMatrix.hpp"
#include "qclab/dense/kron.hpp"
#include "qclab/dense/transpose.hpp"
#include "qclab/io/QASMFile.hpp"
#include "qclab/io/util.hpp"
#include "qclab/qasm.hpp"
#include "qclab/qgates/CNOT.hpp"
#include "qclab/qgates/CPhase.hpp"
#include "qclab/qgates/CRotationX.hpp"
#include "qclab/qgates/CRotationY.hpp"
#include "qclab/qgates/CRotationZ.hpp"
#include "qclab/qgates/CX.hpp"
#include "qclab/qgates/CY.hpp"
#include "qclab/qgates/CZ.hpp"
#include "qclab/qgates/Hadamard.hpp"
#include "qclab/qgates/Identity.hpp"
#include "qclab/qgates/PauliX.hpp"
#include "qclab/qgates/PauliY.hpp"
#include "qclab/qgates/PauliZ.hpp"
#include "qclab/qgates/Phase.hpp"
#include "qclab/qgates/Phase45.hpp"
#include "qclab/qgates/Phase90.hpp"
#include "qclab/qgates/PointerGate1.hpp"
#include "qclab/qgates/PointerGate2.hpp"
#include "qclab/qgates/QControlledGate2.hpp"
#include "qclab/qgates/QGate1.hpp"
#include "qclab/qgates/QGate2.hpp"
#include "qclab/qgates/QRotationGate1.hpp"
#include "qclab/qgates/QRotationGate2.hpp"
#include "qclab/qgates/RotationX.hpp"
#include "qclab/qgates/RotationXX.hpp"
#include "qclab/qgates/RotationY.hpp"
#include "qclab/qgates/RotationYY.hpp"
#include "qclab/qgates/RotationZ.hpp"
#include "qclab/qgates/RotationZZ.hpp"
#include "qclab/qgates/SWAP.hpp"
#include "qclab/qgates/iSWAP.hpp"

// instantiate the class templates for both state types
template class qclab::QCircuit<std::complex<double>>;
template class qclab::QCircuit<std::complex<float>>;
template class qclab::qgates::CPhase<std::complex<double>>;
template class qclab::qgates::CRotationY<std::complex<float>>;
template class qclab::qgates::RotationYY<std::complex<double>>;
template class qclab::qgates::PointerGate2<std::complex<float>>;

int main() { return 0; }

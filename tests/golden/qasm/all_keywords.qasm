OPENQASM 2.0;
include "qelib1.inc";
// every keyword of the reference reader
qreg q[5];
creg c[5];
h q[0];
x q[1]; y q[2];
z q[3];
rx(0.5) q[4];
ry(-1.25) q[0];
rz(3.0e-2) q[1];
p(2.5) q[2];
cx q[0], q[1];
cy q[2],q[3];
cz q[4], q[0];
swap q[1], q[3];
barrier q;
measure q[0] -> c[0];

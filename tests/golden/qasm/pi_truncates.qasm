OPENQASM 2.0;
include "qelib1.inc";
qreg q[3];
h q[0];
rx(0.25) q[1];
cx q[0], q[2];
rz(pi/4) q[1];
h q[2];
x q[0];

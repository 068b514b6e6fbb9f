OPENQASM 2.0;
qreg q[3];
qreg r[2];
h q[0];
h r[1];
x q[2];
cx q[1], q[2];

OPENQASM 2.0;
qreg q[3];
h q[0] x q[1];
ry(0.5) q[2];
cx q[0], q[1] cz q[1], q[2];
z q[2];

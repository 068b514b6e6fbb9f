OPENQASM 2.0;
qreg q[3];
rx(0.1) q[0]; ry(-pi) q[1]; h q[2];
x q[1];

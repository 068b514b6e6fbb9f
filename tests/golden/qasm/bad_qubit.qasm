OPENQASM 2.0;
qreg q[2];
h q[0];
x q[a];
h q[1];

OPENQASM 2.0;
qreg q[2];
	h q[0];
  cx q[0], q[1];

ry(0.125) q[1];

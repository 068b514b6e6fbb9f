OPENQASM 2.0; qreg q[2]; h q[0]; cx q[0],q[1]; rz(0.5) q[1]; swap q[0],q[1];

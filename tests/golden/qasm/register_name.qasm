OPENQASM 2.0;
include "qelib1.inc";
qreg qr[4];
h qr[0];
rx(1.5) qr[3];
cz qr[1], qr[2];
swap qr[0],qr[3];

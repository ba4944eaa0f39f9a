// Direct access, tet loop: the node indices as floats (TetVer read, TetDat
// written).
TetDat = (float4){TetVer.s0, TetVer.s1, TetVer.s2, TetVer.s3};

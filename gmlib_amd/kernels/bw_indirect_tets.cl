// Indirect access, tet loop: sum of the four vertex coordinates read
// through the element -> vertex downlink.
TetDat = VerCrd[0] + VerCrd[1] + VerCrd[2] + VerCrd[3];

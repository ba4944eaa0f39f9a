// Barycentre of a tetrahedron (GmlTetrahedra loop, VerCrd downlink read,
// Bary written).
Bary = (((VerCrd[0] + VerCrd[1]) + VerCrd[2]) + VerCrd[3]) * 0.25f;

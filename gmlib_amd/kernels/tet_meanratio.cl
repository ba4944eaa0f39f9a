// Mean-ratio quality of a tetrahedron (GmlTetrahedra loop, VerCrd read through
// the element->vertex downlink, qal written): 12 cbrt(9 V^2) / sum of the six
// squared edge lengths, signed by the volume.  1 for a regular tet.
{
   float4 e0 = VerCrd[1] - VerCrd[0], e1 = VerCrd[2] - VerCrd[0], e2 = VerCrd[3] - VerCrd[0];
   float4 e3 = VerCrd[2] - VerCrd[1], e4 = VerCrd[3] - VerCrd[1], e5 = VerCrd[3] - VerCrd[2];
   float  s  = dot(e0, e0) + dot(e1, e1) + dot(e2, e2) + dot(e3, e3) + dot(e4, e4) + dot(e5, e5);
   float  v  = dot(cross(e0, e1), e2) * (1.f / 6.f);
   qal = 12.f * cbrt(9.f * v * v) / s;
   if(v < 0.f)
      qal = -qal;
}

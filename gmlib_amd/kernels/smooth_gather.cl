// Jacobi Laplacian step of a vertex towards the mean barycentre of its ball
// (GmlVertices loop, Bary read through the vertex->tet uplink, VerCrd read and
// written with its ref): boundary vertices (ref != 0) stay fixed.
{
   float4 s = (float4)(0.f);
   for(int i = 0; i < BaryDegMax; i++)
      s += Bary[i];
   if(!VerRef && BaryDeg > 0)
   {
      s *= 1.f / (float)BaryDeg;
      VerCrd = VerCrd + (s - VerCrd) * GmlPar->omega;
   }
}

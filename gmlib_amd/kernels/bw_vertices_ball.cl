// Ragged access, vertex loop: sum of TetDat over the vertex ball (the
// vertex -> tet uplink), slots past the degree reading 0.
{
   float4 s = (float4)(0.f);
   for(int i = 0; i < TetDatDegMax; i++)
      s += TetDat[i];
   VerDat = s;
}

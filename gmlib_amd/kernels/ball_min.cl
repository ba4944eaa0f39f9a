// Smallest element value over a vertex ball (GmlVertices loop, qal read through
// the vertex->tet uplink, vq written).
{
   vq = qal[0];
   for(int i = 1; i < qalDeg; i++)
      vq = min(vq, qal[i]);
}

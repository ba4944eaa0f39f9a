/* loop over triangles: the face gradient is the mean of the two cells'
   gradients, the inner cell's for a fixed temperature, the prescribed
   value for a fixed gradient */
GrdExt = 0.5f * (GrdTet[0] + GrdTet[1]);
for(int p = 0; p < 6; p++)
   if(GmlPar->ref[p] == TriRef)
      GrdExt = GmlPar->kind[p] == 0 ? GrdTet[0] : (float4)(GmlPar->val[p]);

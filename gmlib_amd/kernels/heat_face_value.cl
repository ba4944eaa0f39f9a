/* loop over triangles: the face temperature is the mean of the two cells
   sharing it (SolTet[1] is 0 on the boundary), or the patch condition */
SolExt = 0.5f * (SolTet[0] + SolTet[1]);
for(int p = 0; p < 6; p++)
   if(GmlPar->ref[p] == TriRef)
      SolExt = GmlPar->kind[p] == 0 ? GmlPar->val[p] : SolTet[0];

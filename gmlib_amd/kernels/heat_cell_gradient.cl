/* loop over tets: Green-Gauss gradient, sum of the face temperatures times
   the face area vectors (faces in the local order of the tet -> triangle
   link: opposite vertex 0, 1, 2, 3) over the volume */
float4 g = (float4)(0.f);
#define HEAT_FACE(f, a, b, c) g += SolExt[f] * cross(VerCrd[b] - VerCrd[a], VerCrd[c] - VerCrd[a])
HEAT_FACE(0, 1, 2, 3);
HEAT_FACE(1, 2, 0, 3);
HEAT_FACE(2, 3, 0, 1);
HEAT_FACE(3, 0, 2, 1);
#undef HEAT_FACE
GrdTet = 0.5f * g / CalTetVol(VerCrd[0], VerCrd[1], VerCrd[2], VerCrd[3]);

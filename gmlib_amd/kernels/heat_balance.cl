/* loop over tets: flux balance, sum of the face gradients dotted with the
   face area vectors over the volume */
float s = 0.f;
#define HEAT_FLUX(f, a, b, c) s += dot(GrdExt[f], cross(VerCrd[b] - VerCrd[a], VerCrd[c] - VerCrd[a]))
HEAT_FLUX(0, 1, 2, 3);
HEAT_FLUX(1, 2, 0, 3);
HEAT_FLUX(2, 3, 0, 1);
HEAT_FLUX(3, 0, 2, 1);
#undef HEAT_FLUX
Rhs = 0.5f * s / CalTetVol(VerCrd[0], VerCrd[1], VerCrd[2], VerCrd[3]);

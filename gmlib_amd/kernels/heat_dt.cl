/* stable time step of each cell */
dt = 0.3f * pow(CalTetVol(VerCrd[0], VerCrd[1], VerCrd[2], VerCrd[3]), 2.f / 3.f);

/* explicit time step */
SolTet += GmlPar->dt * Rhs;

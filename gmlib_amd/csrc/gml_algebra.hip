// Block-sparse matrices and block vectors of the GMlib C API (placeholder
// entry points until the block SpMV path lands; every call reports that the
// operation is not available).

#include "gml_internal.h"

namespace gml {
struct Matrix { int dummy; };
struct Vector { int dummy; };
}

namespace gml {
void FreeAlgebra(GmlCtx *g)
{
   for(int i = 0; i <= GmlMaxMat; i++) { delete g->mat[i]; g->mat[i] = nullptr; }
   for(int i = 0; i <= GmlMaxVec; i++) { delete g->vec[i]; g->vec[i] = nullptr; }
}
}

using namespace gml;

extern "C" {

static int NotYet(const char *f) { fprintf(stderr, "gml: %s is not available in this build\n", f); return 0; }

int GmlNewMatrix(size_t, int, int, int, void *, int *, int *, int) { return NotYet("GmlNewMatrix"); }
int GmlNewVector(size_t, int, int, void *, int) { return NotYet("GmlNewVector"); }
int GmlMultMatVec(size_t, int, int, int) { NotYet("GmlMultMatVec"); return -1; }
int GmlMultDiagMatVec(size_t, int, int, int) { NotYet("GmlMultDiagMatVec"); return -1; }
int GmlAddVec3(size_t, int, int, int, int) { NotYet("GmlAddVec3"); return -1; }
int GmlScaleVec(size_t, int, double *) { NotYet("GmlScaleVec"); return -2; }
int GmlNormVec(size_t, int, int, double *) { NotYet("GmlNormVec"); return -2; }

}

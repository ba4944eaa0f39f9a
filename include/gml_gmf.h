/* libMeshb keyword and field-type codes used by GmlImportMesh and
 * GmlExportSolution (sources/gmlib3.c:361-373, 4456-4740), for callers that
 * do not include libMeshb's own header (libmeshb8.h), whose values these
 * are.  Dimension, Vertices, Edges, Triangles and End are also the codes
 * found in the reference's sample meshes (the .meshb files of sample_meshes).
 * Include this header or libmeshb8.h, not both. */

#ifndef GML_GMF_H
#define GML_GMF_H

enum gml_gmf_keyword {
   GmfDimension            = 3,
   GmfVertices             = 4,
   GmfEdges                = 5,
   GmfTriangles            = 6,
   GmfQuadrilaterals       = 7,
   GmfTetrahedra           = 8,
   GmfPrisms               = 9,
   GmfHexahedra            = 10,
   GmfSolAtPyramids        = 26,
   GmfPyramids             = 49,
   GmfEnd                  = 54,
   GmfSolAtVertices        = 62,
   GmfSolAtEdges           = 63,
   GmfSolAtTriangles       = 64,
   GmfSolAtQuadrilaterals  = 65,
   GmfSolAtTetrahedra      = 66,
   GmfSolAtPrisms          = 67,
   GmfSolAtHexahedra       = 68
};

/* field types of a solution keyword */
enum gml_gmf_field { GmfSca = 1, GmfVec = 2, GmfSymMat = 3, GmfMat = 4 };

#endif

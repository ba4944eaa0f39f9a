/* Cell-centered heat equation: boundary patches and time step.  A patch
   is the set of boundary triangles of one ref; its condition is a fixed
   temperature (kind 0) or a fixed normal gradient (kind 1). */
typedef struct {
   int   kind[6];
   int   ref[6];
   float val[6];
   float dt;
} GmlParSct;

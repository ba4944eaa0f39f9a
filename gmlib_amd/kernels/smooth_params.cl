typedef struct {
   float omega;
   int   pad[3];
}GmlParSct;

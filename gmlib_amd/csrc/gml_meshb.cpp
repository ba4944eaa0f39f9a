// Mesh and solution files of the GMlib C API: GmlImportMesh and
// GmlExportSolution (sources/gmlib3.c:4456-4740), which the reference
// implements over libMeshb 8 (an external library, absent here).  This file
// reads and writes the same file formats itself:
//
//   .meshb / .solb  binary: int32 code 1 (byte order), int32 version, then
//                   keyword blocks: int32 keyword code, position of the next
//                   keyword (int32 for versions 1-2, int64 from 3), the
//                   line count (int32, int64 from version 4) and the lines.
//                   Reals are float32 in version 1, float64 from 2; node
//                   indices int32, int64 from version 4.  A file ends with
//                   the End keyword.
//   .mesh / .sol    ASCII: keyword names followed by the count and the
//                   lines, '#' comments.
//
// Keyword codes are libMeshb's (include/gml_gmf.h).  Element nodes are
// 1-based in the files and 0-based in the engine.

#include "gml_internal.h"
#include "gml_gmf.h"

#include <sys/types.h>

#include <cctype>
#include <climits>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace gml;

namespace {

// engine mesh kind of a libMeshb element keyword (-1: none)
int Gmf2Gml(int kwd)
{
   static const int tab[GmlMaxEleTyp] = {GmfVertices, GmfEdges, GmfTriangles, GmfQuadrilaterals,
                                         GmfTetrahedra, GmfPyramids, GmfPrisms, GmfHexahedra};
   for(int i = 0; i < GmlMaxEleTyp; i++)
      if(kwd == tab[i]) return i;
   return -1;
}

const int GmfSolKwd[GmlMaxEleTyp] = {GmfSolAtVertices, GmfSolAtEdges, GmfSolAtTriangles, GmfSolAtQuadrilaterals,
                                     GmfSolAtTetrahedra, GmfSolAtPyramids, GmfSolAtPrisms, GmfSolAtHexahedra};
const char *GmfMshNam[GmlMaxEleTyp] = {"Vertices", "Edges", "Triangles", "Quadrilaterals",
                                       "Tetrahedra", "Pyramids", "Prisms", "Hexahedra"};
const char *GmfSolNam[GmlMaxEleTyp] = {"SolAtVertices", "SolAtEdges", "SolAtTriangles", "SolAtQuadrilaterals",
                                       "SolAtTetrahedra", "SolAtPyramids", "SolAtPrisms", "SolAtHexahedra"};

bool EndsWith(const std::string &s, const char *suf)
{
   const size_t n = strlen(suf);
   return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// one mesh keyword read from a file: nodes (or coordinates) and refs
struct Block {
   long long n = 0;
   std::vector<float> crd;   // vertices: 3 per line (z = 0 in 2D)
   std::vector<int> nod;     // elements: EleNmbNod per line, 0-based
   std::vector<int> ref;
};

// --- binary reader ---------------------------------------------------------

struct BinFile {
   FILE *f = nullptr;
   bool swap = false;
   int ver = 0, dim = 0;
   long long size = 0;         // file size in bytes
   ~BinFile() { if(f) fclose(f); }

   bool Read(void *p, size_t siz, size_t cnt)
   {
      if(fread(p, siz, cnt, f) != cnt) return false;
      if(swap && siz > 1)
      {
         unsigned char *b = (unsigned char *)p;
         for(size_t i = 0; i < cnt; i++, b += siz)
            for(size_t k = 0; k < siz / 2; k++) std::swap(b[k], b[siz - 1 - k]);
      }
      return true;
   }
   bool Int(long long *v, bool wide)
   {
      if(wide) { int64_t x; if(!Read(&x, 8, 1)) return false; *v = x; }
      else     { int32_t x; if(!Read(&x, 4, 1)) return false; *v = x; }
      return true;
   }
   bool Real(float *v)
   {
      if(ver >= 2) { double x; if(!Read(&x, 8, 1)) return false; *v = (float)x; }
      else         { float x; if(!Read(&x, 4, 1)) return false; *v = x; }
      return true;
   }
};

// keyword positions of a binary file: code -> (data position, line count)
struct BinKwd { int kwd; long long pos, n; };

// smallest size in bytes of one line of mesh kind typ in a binary file
long long MinLine(const BinFile &b, int typ)
{
   const long long isz = b.ver >= 4 ? 8 : 4, rsz = b.ver >= 2 ? 8 : 4;
   return typ == GmlVertices ? b.dim * rsz + isz : (EleNmbNod[typ] + 1) * isz;
}

// The file is not trusted: a keyword's next position must lie past what was
// read of it and inside the file (so the scan always advances), its line
// count must be >= 0 and its lines must fit in the rest of the file.
bool ScanBinary(BinFile &b, std::vector<BinKwd> *kw)
{
   for(;;)
   {
      long long kwd = 0, nxt = 0, n = 0;
      const long long start = (long long)ftello(b.f);
      if(!b.Int(&kwd, false)) return false;
      if(kwd == GmfEnd) return true;
      if(!b.Int(&nxt, b.ver >= 3)) return false;
      if(kwd == GmfDimension)
      {
         long long d = 0;
         if(!b.Int(&d, false)) return false;
         if(d < 1 || d > 3) return false;
         b.dim = (int)d;
      }
      else if(Gmf2Gml((int)kwd) >= 0)
      {
         if(!b.Int(&n, b.ver >= 4)) return false;
         const long long pos = (long long)ftello(b.f);
         if(n < 0 || n > (b.size - pos) / MinLine(b, Gmf2Gml((int)kwd))) return false;
         kw->push_back({(int)kwd, pos, n});
      }
      if(!nxt) return true;
      if(nxt <= start || nxt < (long long)ftello(b.f) || nxt > b.size) return false;
      if(fseeko(b.f, (off_t)nxt, SEEK_SET)) return false;
   }
}

bool ReadBinaryBlock(BinFile &b, const BinKwd &k, int typ, Block *blk)
{
   if(fseeko(b.f, (off_t)k.pos, SEEK_SET)) return false;
   blk->n = k.n;
   blk->ref.resize(k.n);
   if(typ == GmlVertices)
   {
      blk->crd.assign(k.n * 3, 0.f);
      for(long long i = 0; i < k.n; i++)
      {
         for(int j = 0; j < b.dim; j++)
         {
            float x;
            if(!b.Real(&x)) return false;
            if(j < 3) blk->crd[i * 3 + j] = x;
         }
         long long r;
         if(!b.Int(&r, b.ver >= 4)) return false;
         blk->ref[i] = (int)r;
      }
   }
   else
   {
      const int nn = EleNmbNod[typ];
      blk->nod.resize(k.n * nn);
      for(long long i = 0; i < k.n; i++)
      {
         for(int j = 0; j < nn; j++)
         {
            long long v;
            if(!b.Int(&v, b.ver >= 4)) return false;
            blk->nod[i * nn + j] = (int)(v - 1);
         }
         long long r;
         if(!b.Int(&r, b.ver >= 4)) return false;
         blk->ref[i] = (int)r;
      }
   }
   return true;
}

// --- ASCII reader ----------------------------------------------------------

struct Text {
   std::vector<char> buf;
   size_t at = 0;
   // next token (skips blanks and '#' comments); empty at the end
   std::string Tok()
   {
      for(;;)
      {
         while(at < buf.size() && isspace((unsigned char)buf[at])) at++;
         if(at < buf.size() && buf[at] == '#') { while(at < buf.size() && buf[at] != '\n') at++; continue; }
         break;
      }
      const size_t b = at;
      while(at < buf.size() && !isspace((unsigned char)buf[at])) at++;
      return at > b ? std::string(buf.data() + b, at - b) : std::string();
   }
   bool IsNumber(const std::string &t) const
   {
      return !t.empty() && (isdigit((unsigned char)t[0]) || t[0] == '-' || t[0] == '+' || t[0] == '.');
   }
};

bool ReadAscii(const char *nam, const std::vector<int> &want, std::vector<Block> *out, std::vector<bool> *found)
{
   FILE *f = fopen(nam, "rb");
   if(!f) return false;
   Text t;
   fseeko(f, 0, SEEK_END);
   t.buf.resize((size_t)ftello(f));
   fseeko(f, 0, SEEK_SET);
   const bool ok = fread(t.buf.data(), 1, t.buf.size(), f) == t.buf.size();
   fclose(f);
   if(!ok) return false;
   int dim = 3;
   std::string tok = t.Tok();
   while(!tok.empty() && tok != "End")
   {
      int typ = -1;
      for(int i = 0; i < GmlMaxEleTyp; i++)
         if(tok == GmfMshNam[i]) typ = i;
      if(tok == "MeshVersionFormatted") { t.Tok(); tok = t.Tok(); continue; }
      if(tok == "Dimension")
      {
         dim = atoi(t.Tok().c_str());
         if(dim < 1 || dim > 3) return false;
         tok = t.Tok();
         continue;
      }
      int slot = -1;
      for(size_t k = 0; k < want.size(); k++)
         if(want[k] == typ) slot = (int)k;
      if(typ < 0 || slot < 0)
      {
         // another keyword: skip its numbers
         tok = t.Tok();
         while(t.IsNumber(tok)) tok = t.Tok();
         continue;
      }
      Block &b = (*out)[slot];
      b.n = atoll(t.Tok().c_str());
      // every line takes at least two characters per number
      if(b.n < 0 || b.n > (long long)(t.buf.size() / 2)) return false;
      b.ref.resize(b.n);
      if(typ == GmlVertices)
      {
         b.crd.assign(b.n * 3, 0.f);
         for(long long i = 0; i < b.n; i++)
         {
            for(int j = 0; j < dim; j++)
            {
               const float x = strtof(t.Tok().c_str(), nullptr);
               if(j < 3) b.crd[i * 3 + j] = x;
            }
            b.ref[i] = atoi(t.Tok().c_str());
         }
      }
      else
      {
         const int nn = EleNmbNod[typ];
         b.nod.resize(b.n * nn);
         for(long long i = 0; i < b.n; i++)
         {
            for(int j = 0; j < nn; j++) b.nod[i * nn + j] = (int)(atoll(t.Tok().c_str()) - 1);
            b.ref[i] = atoi(t.Tok().c_str());
         }
      }
      (*found)[slot] = true;
      tok = t.Tok();
   }
   return true;
}

// new mesh table of kind typ holding a block (host mirror filled, uploaded);
// element node indices must lie in the vertex table (nver vertices)
int Install(size_t GmlIdx, GmlCtx *g, int typ, const Block &b, long long nver)
{
   if(b.n < 1 || b.n >= INT_MAX) return 0;
   if(typ != GmlVertices)
      for(int v : b.nod)
         if(v < 0 || v >= nver)
         {
            printf("%s: a node index is outside the %lld vertices\n", GmfMshNam[typ], nver);
            return 0;
         }
   const int idx = GmlNewMeshData(GmlIdx, typ, (int)b.n);
   if(!idx) return 0;
   Data *d = &g->dat[idx], *r = &g->dat[g->RefIdx[typ]];
   if(!EnsureHost(d) || !EnsureHost(r)) return 0;
   if(typ == GmlVertices)
   {
      float *c = (float *)d->host;
      for(long long i = 0; i < b.n; i++)
      {
         for(int j = 0; j < 3; j++) c[i * 4 + j] = b.crd[i * 3 + j];
         c[i * 4 + 3] = 0.f;
      }
   }
   else
   {
      const int w = d->ItmLen, nn = EleNmbNod[typ];
      int *e = (int *)d->host;
      memset(e, 0, d->MemSiz);
      for(long long i = 0; i < b.n; i++)
         for(int j = 0; j < nn; j++) e[i * w + j] = b.nod[i * nn + j];
   }
   memcpy(r->host, b.ref.data(), (size_t)b.n * 4);
   d->host_valid = r->host_valid = true;
   if(!Upload(g, idx) || !Upload(g, g->RefIdx[typ])) return 0;
   return idx;
}

// --- solution writer -------------------------------------------------------

struct Field { int idx; int nsca; };   // a data item and its scalars per line

double Scalar(const Data *d, long long lin, int k)
{
   const char *p = (const char *)d->host + (size_t)lin * d->LinSiz;
   const int t = d->ItmTyp;
   if(t >= GmlDbl && t <= GmlDbl16) return ((const double *)p)[k];
   if(t >= GmlFlt && t <= GmlFlt16) return ((const float *)p)[k];
   if(t >= GmlByt && t <= GmlByt16) return ((const signed char *)p)[k];
   return ((const int *)p)[k];
}

} // namespace

extern "C" {

// Read the requested element keywords (a 0-terminated list of libMeshb
// codes, at most 10) of a .mesh / .meshb file into new mesh tables.
// Returns the number of requested keywords the engine knows (present in the
// file or not), 0 if the file cannot be read.
static int ImportMesh(size_t GmlIdx, const char *MshNam, const std::vector<int> &want);

int GmlImportMesh(size_t GmlIdx, char *MshNam, ...)
{
   std::vector<int> want;
   va_list ap;
   va_start(ap, MshNam);
   for(int kwd; (kwd = va_arg(ap, int)) && want.size() < 10;)
   {
      const int typ = Gmf2Gml(kwd);
      if(typ >= 0) want.push_back(typ);
   }
   va_end(ap);
   if(!MshNam) return 0;
   // nothing thrown by the reader (an allocation failure) crosses the C ABI
   try
   {
      return ImportMesh(GmlIdx, MshNam, want);
   }
   catch(...)
   {
      printf("%s: cannot be read\n", MshNam);
      return 0;
   }
}

static int ImportMesh(size_t GmlIdx, const char *MshNam, const std::vector<int> &want)
{
   GML_CTX(g, GmlIdx);
   const std::string nam = MshNam;
   std::vector<Block> blk(want.size());
   std::vector<bool> found(want.size(), false);
   if(EndsWith(nam, ".mesh"))
   {
      if(!ReadAscii(MshNam, want, &blk, &found))
      {
         printf("Could not read file %s\n", MshNam);
         return 0;
      }
   }
   else
   {
      BinFile b;
      int32_t code = 0, ver = 0;
      std::vector<BinKwd> kw;
      if(!(b.f = fopen(MshNam, "rb")) || fseeko(b.f, 0, SEEK_END) || (b.size = (long long)ftello(b.f)) < 8
      || fseeko(b.f, 0, SEEK_SET) || fread(&code, 4, 1, b.f) != 1)
      {
         printf("Could not open file %s\n", MshNam);
         return 0;
      }
      if(code == 16777216) b.swap = true;
      else if(code != 1)
      {
         printf("%s is not a libMeshb binary file\n", MshNam);
         return 0;
      }
      if(!b.Read(&ver, 4, 1) || ver < 1 || ver > 4)
      {
         printf("%s: unsupported version %d\n", MshNam, ver);
         return 0;
      }
      b.ver = ver;
      b.dim = 3;
      if(!ScanBinary(b, &kw))
      {
         printf("%s: truncated or inconsistent file\n", MshNam);
         return 0;
      }
      for(size_t k = 0; k < want.size(); k++)
         for(const BinKwd &x : kw)
            if(Gmf2Gml(x.kwd) == want[k])
            {
               if(!ReadBinaryBlock(b, x, want[k], &blk[k]))
               {
                  printf("%s: truncated keyword %d\n", MshNam, x.kwd);
                  return 0;
               }
               found[k] = true;
            }
   }
   // vertices first: element kinds inherit nothing from them, but a handle
   // is usually filled in that order
   for(int pass = 0; pass < 2; pass++)
      for(size_t k = 0; k < want.size(); k++)
         if(found[k] && (want[k] == GmlVertices) == (pass == 0) && !g->TypIdx[want[k]])
            Install(GmlIdx, g, want[k], blk[k], g->NmbEle[GmlVertices]);
   return (int)want.size();
}

// Write data items (a 0-terminated list of data indices, at most 10) to a
// .sol / .solb file: one SolAt keyword per mesh kind, whose lines hold every
// scalar of every listed item of that kind in list order.  Returns the
// number of keywords written, 0 on failure.
int GmlExportSolution(size_t GmlIdx, char *SolNam, ...)
{
   GML_CTX(g, GmlIdx);
   std::vector<Field> fld[GmlMaxEleTyp];
   std::vector<int> order;
   va_list ap;
   va_start(ap, SolNam);
   int cnt = 0;
   for(int idx; (idx = va_arg(ap, int)) && cnt < 10; cnt++)
   {
      if(idx < 1 || idx > GmlMaxDat || !g->dat[idx].use) continue;
      Data *d = &g->dat[idx];
      if(d->kind != DkRaw && d->kind != DkEle) continue;
      const int typ = d->MshTyp;
      if(typ < 0 || typ >= GmlMaxEleTyp) continue;
      if(!d->host_valid && !Download(g, idx)) continue;
      if(fld[typ].empty()) order.push_back(typ);
      const int nsca = d->kind == DkEle ? (typ == GmlVertices ? 3 : EleNmbNod[typ]) : d->NmbItm * d->ItmLen;
      fld[typ].push_back({idx, nsca});
   }
   va_end(ap);
   if(!SolNam) return 0;
   const std::string nam = SolNam;
   const bool ascii = EndsWith(nam, ".sol");
   FILE *f = fopen(SolNam, ascii ? "w" : "wb");
   if(!f)
   {
      printf("Could not create file %s\n", SolNam);
      return 0;
   }
   int nkwd = 0;
   if(ascii)
   {
      fprintf(f, "MeshVersionFormatted 1\n\nDimension 3\n\n");
      for(int typ : order)
      {
         const long long n = g->NmbEle[typ];
         int ns = 0;
         for(const Field &x : fld[typ]) ns += x.nsca;
         fprintf(f, "%s\n%lld\n%d", GmfSolNam[typ], n, ns);
         for(int k = 0; k < ns; k++) fprintf(f, " %d", GmfSca);
         fprintf(f, "\n");
         for(long long i = 0; i < n; i++)
         {
            const char *sep = "";
            for(const Field &x : fld[typ])
               for(int k = 0; k < x.nsca; k++)
               {
                  fprintf(f, "%s%.9g", sep, Scalar(&g->dat[x.idx], i, k));
                  sep = " ";
               }
            fprintf(f, "\n");
         }
         fprintf(f, "\n");
         nkwd++;
      }
      // names of the fields (ReferenceStrings: solution keyword, index, name)
      int nstr = 0;
      for(int typ : order)
         for(const Field &x : fld[typ]) nstr += x.nsca;
      fprintf(f, "ReferenceStrings\n%d\n", nstr);
      static const char DimChr[4] = {'x', 'y', 'z', 't'};
      int c = 0;
      for(int typ : order)
         for(const Field &x : fld[typ])
            for(int k = 0; k < x.nsca; k++)
            {
               c++;
               const std::string &dn = g->dat[x.idx].nam;
               if(x.nsca > 1 && k < 4) fprintf(f, "%d 1 %s.%c %d\n", GmfSolKwd[typ], dn.c_str(), DimChr[k], c);
               else                   fprintf(f, "%d 1 %s %d\n", GmfSolKwd[typ], dn.c_str(), c);
            }
      fprintf(f, "\nEnd\n");
   }
   else
   {
      // version 1 (float32 reals, int32 positions), dimension 3
      auto put32 = [&](int32_t v) { fwrite(&v, 4, 1, f); };
      put32(1);
      put32(1);
      put32(GmfDimension);
      put32((int32_t)(ftello(f) + 8));
      put32(3);
      for(int typ : order)
      {
         const long long n = g->NmbEle[typ];
         int ns = 0;
         for(const Field &x : fld[typ]) ns += x.nsca;
         const long long next = ftello(f) + 4 + 4 + 4 + 4 + 4LL * ns + 4LL * ns * n;
         if(next >= INT32_MAX) { fclose(f); printf("%s: too large for a version 1 file\n", SolNam); return 0; }
         put32(GmfSolKwd[typ]);
         put32((int32_t)next);
         put32((int32_t)n);
         put32(ns);
         for(int k = 0; k < ns; k++) put32(GmfSca);
         std::vector<float> line(ns);
         for(long long i = 0; i < n; i++)
         {
            int at = 0;
            for(const Field &x : fld[typ])
               for(int k = 0; k < x.nsca; k++) line[at++] = (float)Scalar(&g->dat[x.idx], i, k);
            fwrite(line.data(), 4, ns, f);
         }
         nkwd++;
      }
      put32(GmfEnd);
      put32(0);
   }
   const bool ok = !ferror(f);
   fclose(f);
   return ok ? nkwd : 0;
}

} // extern "C"

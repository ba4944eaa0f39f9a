// Kernel engine of libgml: GmlCompileKernel (argument classification, HIP
// source generation around the user's OpenCL-C body, hipRTC for gfx950,
// code-object cache) and GmlLaunchKernel (launch-shape tuner, profiling).
//
// Reference behaviour followed (paths in the GMlib tree):
//   GmlCompileKernel argument classification  sources/gmlib3.c:2217-2499
//   generated-kernel contract (names, locals,
//   counters, reads, body, writes)            sources/gmlib3.c:2502-2938
//   high-degree twin kernel                   sources/gmlib3.c:2530-2577
//   NewOclKrn compile options                 sources/gmlib3.c:2971-3057
//   GmlLaunchKernel / RunOclKrn               sources/gmlib3.c:3064-3222
//   GmlGetKernelRunTime                       sources/gmlib3.c:4343-4366
//
// What differs by design: uplink rows are read from a SELL-64 layout (one
// coalesced 256-byte load per ball slot and wave) instead of per-lane int16
// vectors; uplink data of several items per line are fetched lazily, only
// when the body reads them; launches are asynchronous on the handle's
// stream; the work-group size (and XCD block remap) is tuned with HIP events
// without serialising the queue.

#include "gml_internal.h"

#include <mutex>

#include <hip/hiprtc.h>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace gml {

static const char *kPrelude =
#include "gml_prelude.inc"
;

// ---------------------------------------------------------------------------
// -cl-single-precision-constant: unsuffixed floating literals of user code
// become float literals.  Comments, strings and character literals are
// skipped; hexadecimal literals are left alone.
// ---------------------------------------------------------------------------
static std::string SinglePrecisionLiterals(const char *s)
{
   std::string out;
   if(!s) return out;
   size_t n = strlen(s), i = 0;
   out.reserve(n + n / 16);
   auto isid = [](char c) { return isalnum((unsigned char)c) || c == '_'; };

   while(i < n)
   {
      char c = s[i];
      if(c == '/' && i + 1 < n && s[i + 1] == '/')
      {
         size_t j = i;
         while(j < n && s[j] != '\n') j++;
         out.append(s + i, j - i); i = j; continue;
      }
      if(c == '/' && i + 1 < n && s[i + 1] == '*')
      {
         size_t j = i + 2;
         while(j + 1 < n && !(s[j] == '*' && s[j + 1] == '/')) j++;
         j = j + 2 <= n ? j + 2 : n;
         out.append(s + i, j - i); i = j; continue;
      }
      if(c == '"' || c == '\'')
      {
         size_t j = i + 1;
         while(j < n && s[j] != c) { if(s[j] == '\\') j++; j++; }
         j = j < n ? j + 1 : n;
         out.append(s + i, j - i); i = j; continue;
      }
      if(isid(c) && !isdigit((unsigned char)c))
      {
         size_t j = i;
         while(j < n && isid(s[j])) j++;
         out.append(s + i, j - i); i = j; continue;
      }
      if(isdigit((unsigned char)c) || (c == '.' && i + 1 < n && isdigit((unsigned char)s[i + 1])))
      {
         // pp-number: digits, letters, '.', and exponent signs
         size_t j = i;
         bool hex = (c == '0' && i + 1 < n && (s[i + 1] == 'x' || s[i + 1] == 'X'));
         while(j < n)
         {
            char d = s[j];
            if(isalnum((unsigned char)d) || d == '.' || d == '_') { j++; continue; }
            if((d == '+' || d == '-') && j > i && (s[j - 1] == 'e' || s[j - 1] == 'E' || (hex && (s[j - 1] == 'p' || s[j - 1] == 'P'))))
            { j++; continue; }
            break;
         }
         std::string tok(s + i, j - i);
         bool flt = false;
         if(!hex)
            for(char d : tok) if(d == '.' || d == 'e' || d == 'E') { flt = true; break; }
         if(flt)
         {
            char last = tok.back();
            if(last == 'l' || last == 'L') tok.pop_back();
            last = tok.back();
            if(last != 'f' && last != 'F') tok.push_back('f');
         }
         out += tok; i = j; continue;
      }
      out.push_back(c); i++;
   }
   return out;
}

// ---------------------------------------------------------------------------
// OpenCL vector literals of several parts, (float4)(a, b, c, d) or
// (float4)(v.xy, w.zw): a C++ compiler reads them as the cast of a comma
// expression (every component = the last part), so they are rewritten to
// gml_vmake<float4>(...), the prelude's component-wise constructor.  A
// single part, (float4)(x), is a splat in both languages and stays.  Parts
// are rewritten recursively; comments, strings and character literals are
// copied as they are.
// ---------------------------------------------------------------------------
static bool IsVectorType(const std::string &t)
{
   static const char *base[] = {"char", "uchar", "short", "ushort", "int", "uint", "long", "ulong", "float", "double"};
   for(const char *b : base)
   {
      const size_t n = strlen(b);
      if(t.size() > n && t.compare(0, n, b) == 0)
      {
         const std::string w = t.substr(n);
         if(w == "2" || w == "3" || w == "4" || w == "8" || w == "16") return true;
      }
   }
   return false;
}

// end of the comment, string or character literal starting at i (i itself
// if none starts there)
static size_t SkipOpaque(const std::string &s, size_t i)
{
   const size_t n = s.size();
   if(s[i] == '/' && i + 1 < n && s[i + 1] == '/')
   {
      while(i < n && s[i] != '\n') i++;
      return i;
   }
   if(s[i] == '/' && i + 1 < n && s[i + 1] == '*')
   {
      size_t j = i + 2;
      while(j + 1 < n && !(s[j] == '*' && s[j + 1] == '/')) j++;
      return j + 2 <= n ? j + 2 : n;
   }
   if(s[i] == '"' || s[i] == '\'')
   {
      const char q = s[i];
      size_t j = i + 1;
      while(j < n && s[j] != q) { if(s[j] == '\\') j++; j++; }
      return j < n ? j + 1 : n;
   }
   return i;
}

static std::string VectorLiterals(const std::string &s)
{
   std::string out;
   const size_t n = s.size();
   size_t i = 0;
   auto isid = [](char c) { return isalnum((unsigned char)c) || c == '_'; };
   auto ws = [&](size_t j) { while(j < n && isspace((unsigned char)s[j])) j++; return j; };
   while(i < n)
   {
      const size_t o = SkipOpaque(s, i);
      if(o != i) { out.append(s, i, o - i); i = o; continue; }
      if(s[i] == '(' && (i == 0 || !isid(s[i - 1])))
      {
         // "(" type ")" "(" parts ")"
         size_t j = ws(i + 1), t0 = j;
         while(j < n && isid(s[j])) j++;
         const std::string typ = s.substr(t0, j - t0);
         size_t k = ws(j);
         if(IsVectorType(typ) && k < n && s[k] == ')')
         {
            size_t a = ws(k + 1);
            if(a < n && s[a] == '(')
            {
               // matching parenthesis and top-level commas of the parts
               int depth = 0, commas = 0;
               size_t e = a;
               for(; e < n; e++)
               {
                  const size_t q = SkipOpaque(s, e);
                  if(q != e) { e = q - 1; continue; }
                  const char c = s[e];
                  if(c == '(' || c == '[' || c == '{') depth++;
                  else if(c == ')' || c == ']' || c == '}') { if(--depth == 0) break; }
                  else if(c == ',' && depth == 1) commas++;
               }
               if(e < n && commas > 0)
               {
                  out += "gml_vmake<" + typ + ">(" + VectorLiterals(s.substr(a + 1, e - a - 1)) + ")";
                  i = e + 1;
                  continue;
               }
            }
         }
      }
      out.push_back(s[i]);
      i++;
   }
   return out;
}

static std::string UserCode(const char *s)
{
   return VectorLiterals(SinglePrecisionLiterals(s));
}

// ---------------------------------------------------------------------------
// Code-object cache shared by every handle of the process (hipRTC output for
// gfx950 keyed by a 64-bit FNV-1a hash of source and options).
// ---------------------------------------------------------------------------
static std::mutex g_cache_mtx;
static std::unordered_map<uint64_t, std::string> g_cache;

static uint64_t Fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull)
{
   for(unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
   return h;
}

static bool RtcCompile(const std::string &src, const std::string &name, const std::string &cflags,
                       std::string *code, std::string *log)
{
   std::vector<std::string> opts = {
      "--gpu-architecture=gfx950", "-O3", "-ffp-contract=fast", "-std=c++17",
      "-Wno-c++11-narrowing", "-Wno-unused-value", "-Wno-unused-variable",
      "-Wno-macro-redefined", "-Wno-unused-but-set-variable",
      "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
   // user options, split on blanks (GmlSetCompilerOptions)
   {
      size_t i = 0;
      while(i < cflags.size())
      {
         while(i < cflags.size() && isspace((unsigned char)cflags[i])) i++;
         size_t j = i;
         while(j < cflags.size() && !isspace((unsigned char)cflags[j])) j++;
         if(j > i)
         {
            std::string o = cflags.substr(i, j - i);
            if(o.rfind("-cl-", 0) != 0)          // OpenCL-only options have no meaning here
               opts.push_back(o);
         }
         i = j;
      }
   }
   std::string key = src;
   for(auto &o : opts) { key += '\x01'; key += o; }
   uint64_t h = Fnv1a(key);
   {
      std::lock_guard<std::mutex> lk(g_cache_mtx);
      auto it = g_cache.find(h);
      if(it != g_cache.end()) { *code = it->second; return true; }
   }
   hiprtcProgram prog;
   if(hiprtcCreateProgram(&prog, src.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
   {
      *log = "hiprtcCreateProgram failed";
      return false;
   }
   std::vector<const char *> o;
   for(auto &x : opts) o.push_back(x.c_str());
   hiprtcResult r = hiprtcCompileProgram(prog, (int)o.size(), o.data());
   size_t ls = 0;
   hiprtcGetProgramLogSize(prog, &ls);
   if(ls > 1)
   {
      std::string l(ls, '\0');
      hiprtcGetProgramLog(prog, &l[0]);
      *log = l;
   }
   if(r != HIPRTC_SUCCESS)
   {
      hiprtcDestroyProgram(&prog);
      return false;
   }
   size_t cs = 0;
   hiprtcGetCodeSize(prog, &cs);
   code->assign(cs, '\0');
   hiprtcGetCode(prog, &(*code)[0]);
   hiprtcDestroyProgram(&prog);
   std::lock_guard<std::mutex> lk(g_cache_mtx);
   g_cache[h] = *code;
   return true;
}

// ---------------------------------------------------------------------------
// Argument descriptors (the reference's ArgSct, gmlib3.c:86-91)
// ---------------------------------------------------------------------------
enum ArgKind { AkData = 0, AkBall, AkDeg, AkDown, AkTrans };

struct Arg {
   int kind = AkData;
   int MshTyp = 0, DatIdx = 0, LnkDir = 0, LnkIdx = -1, CntIdx = -1;
   int LnkDeg = -1, MaxDeg = -1, NmbItm = 1, ItmLen = 1, ItmTyp = GmlInt, FlgTab = 0;
   std::string nam, VoyNam;
   int SrcTyp = 0, DstTyp = 0;   // for AkBall: which uplink
};

static std::string Fmt(const char *fmt, ...)
{
   char buf[2048];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   return buf;
}

static const char *HexDig = "0123456789abcdef";

// Pointer type of an argument in the kernel signature.
static std::string ArgDecl(const Arg &a)
{
   if(a.kind == AkBall)
      return Fmt("const int *__restrict__ %sTab, const unsigned int *__restrict__ %sOff", a.nam.c_str(), a.nam.c_str());
   if(a.NmbItm <= 1)
      return Fmt("%s *%sTab", OclTypStr[a.ItmTyp], a.nam.c_str());
   return Fmt("%s (*%sTab)[%d]", OclTypStr[a.ItmTyp], a.nam.c_str(), a.NmbItm);
}

// Emit the generated kernel.  lazy = 1 makes multi-item uplink data lazy
// accessors instead of eagerly filled local arrays.
// ---------------------------------------------------------------------------
// LDS-staged uplink data.  Read-only, single-item data reached through a
// vertex ball (or any uplink) are gathered per work-group: the tile's
// distinct destination entities (ascending ids, so the loads coalesce) are
// loaded once into LDS, and every ball slot then reads LDS through its
// tile-local index (StagePlan).  Values and their order are those of the
// direct gather; only where the bytes come from changes.
// ---------------------------------------------------------------------------
// the ball whose data are staged (the first ball with stageable data), or
// -1; *staged receives the staged data arguments
static int StagedArgs(const std::vector<Arg> &args, std::vector<int> *staged)
{
   int ball = -1;
   size_t bytes = 0;
   staged->clear();
   for(size_t i = 0; i < args.size(); i++)
   {
      const Arg &a = args[i];
      if(a.kind != AkData || a.CntIdx < 0 || a.LnkIdx < 0 || args[a.LnkIdx].kind != AkBall) continue;
      if(!(a.FlgTab & GmlReadMode) || (a.FlgTab & (GmlWriteMode | GmlManual)) || a.NmbItm != 1) continue;
      // a short ball (a face's two cells) gains nothing from the tile's
      // shared loads: its rows are gathered directly
      if(args[a.LnkIdx].MaxDeg <= 4) continue;
      if(ball >= 0 && a.LnkIdx != ball) continue;
      // the smallest tile (64 rows) must fit the LDS whatever its sharing:
      // at most 64 x the ball width entities, each item 16-byte aligned.
      // Data that would not fit are gathered directly.
      const size_t z = ((size_t)OclTypSiz[a.ItmTyp] + 15) & ~(size_t)15;
      if((bytes + z) * 64 * args[a.LnkIdx].MaxDeg > 64 * 1024) continue;
      bytes += z;
      ball = a.LnkIdx;
      staged->push_back((int)i);
   }
   return ball;
}

static bool IsIdent(char c) { return isalnum((unsigned char)c) || c == '_'; }

// does the body name this identifier?
static bool BodyUses(const char *body, const std::string &id)
{
   for(const char *p = strstr(body, id.c_str()); p; p = strstr(p + 1, id.c_str()))
      if((p == body || !IsIdent(p[-1])) && !IsIdent(p[id.size()])) return true;
   return false;
}

// A kernel may also be emitted in paired form (pair = true): each thread
// takes two rows, cnt and cnt + blockDim.x, with the reads of both rows
// issued before either body runs, so that twice as many loads are in flight
// per thread.  The row's locals become the members of a local struct; the
// reads, the body and the writes are lambdas that bind the members under
// their usual names, so the emitted reads, the user's body and the writes
// are the same text as in the one-row form.  A body that returns early
// skips its row's writes, as in the one-row form.  In the last work-group
// a thread without a second row runs the body twice on its first row and
// writes once, so bodies that write the parameter block or use atomics keep
// the one-row form.
// true if the body assigns a field of the parameter block (GmlPar->x = ...,
// GmlPar->x[i] += ..., ++GmlPar->x, ...): such a body must run once per row
static bool BodyWritesParams(const char *body)
{
   for(const char *p = strstr(body, "GmlPar->"); p; p = strstr(p + 1, "GmlPar->"))
   {
      const char *b = p;
      while(b > body && (b[-1] == ' ' || b[-1] == '\t')) b--;
      if(b - body >= 2 && ((b[-1] == '+' && b[-2] == '+') || (b[-1] == '-' && b[-2] == '-'))) return true;
      const char *q = p + 8;
      while(IsIdent(*q)) q++;
      while(*q == '[' || *q == ' ' || *q == '\t' || *q == '.')
      {
         if(*q == '[') { int depth = 1; q++; while(*q && depth) { depth += (*q == '[') - (*q == ']'); q++; } }
         else if(*q == '.') { q++; while(IsIdent(*q)) q++; }
         else q++;
      }
      if((q[0] == '+' && q[1] == '+') || (q[0] == '-' && q[1] == '-')) return true;
      if(q[0] == '=' && q[1] != '=') return true;
      if(strchr("+-*/%&|^", q[0]) && q[0] && q[1] == '=') return true;
      if((q[0] == '<' && q[1] == '<' && q[2] == '=') || (q[0] == '>' && q[1] == '>' && q[2] == '=')) return true;
   }
   return false;
}

// ---------------------------------------------------------------------------
// Element-tile form.  An element loop without uplinks that reads vertex data
// through its own element table (VerCrd[k] of a tet kernel) and whose body
// never names the table row can run over tiles of kDownTile elements: the
// tile's distinct vertices (a DownPlan) are loaded once into LDS with
// coalesced loads, and each row reads its slots there through one-byte
// tile-local indices instead of reading its table row and gathering every
// slot from memory.  Values are those of the direct form; only where the
// bytes come from changes.  The form is one more candidate of the launch
// tuner.  Returns the table argument and the staged data arguments, or -1.
// ---------------------------------------------------------------------------
static const char *kRowIds[] = {"barrier", "get_global_id", "get_local_id", "get_group_id", "get_local_size",
                                "get_global_size", "get_num_groups", "__syncthreads", "gml_gid"};

static bool UsesRowIds(const char *KrnSrc)
{
   for(const char *id : kRowIds)
      if(BodyUses(KrnSrc, id)) return true;
   return false;
}

static int TileStagedArgs(const GmlCtx *g, const std::vector<Arg> &args, int MshTyp, const char *KrnSrc,
                          std::vector<int> *staged)
{
   staged->clear();
   if(MshTyp <= GmlVertices || MshTyp >= GmlMaxEleTyp) return -1;
   int L = -1;
   for(size_t i = 0; i < args.size(); i++)
   {
      const Arg &a = args[i];
      if(a.kind == AkBall || a.kind == AkDeg || (a.FlgTab & GmlManual)) return -1;
      if(a.kind == AkDown && a.DatIdx == g->TypIdx[MshTyp] && a.MshTyp == GmlVertices) L = (int)i;
   }
   if(L < 0 || BodyUses(KrnSrc, args[L].nam)) return -1;
   const int w = args[L].NmbItm * args[L].ItmLen;
   if(w < 1 || w > 16) return -1;
   size_t bytes = 0;
   for(size_t i = 0; i < args.size(); i++)
   {
      const Arg &a = args[i];
      if(a.LnkIdx != L) continue;
      if(!(a.FlgTab & GmlReadMode) || (a.FlgTab & GmlWriteMode) || a.NmbItm != 1 || a.CntIdx >= 0) return -1;
      bytes += OclTypSiz[a.ItmTyp];
      staged->push_back((int)i);
   }
   // a tile holds at most 256 vertices: their data must fit 64 KB of LDS
   if(staged->empty() || bytes * 256 > 64 * 1024) return -1;
   for(const char *id : kRowIds)
      if(BodyUses(KrnSrc, id)) return -1;
   return L;
}

static bool PairEligible(const std::vector<Arg> &args, const char *KrnSrc, bool lazy, bool stage)
{
   if(lazy || stage || BodyWritesParams(KrnSrc) || strstr(KrnSrc, "atomic")) return false;
   // uplinks only while a row's gathered values stay small (32 words)
   int words = 0;
   for(const Arg &a : args)
   {
      if(a.FlgTab & GmlManual) return false;
      if(a.kind == AkBall) words += 2 * a.MaxDeg;
      else if(a.CntIdx >= 0) words += args[a.LnkIdx].MaxDeg * a.NmbItm * ((OclTypSiz[a.ItmTyp] + 3) / 4);
   }
   if(words > 32) return false;
   for(const char *id : kRowIds)
      if(BodyUses(KrnSrc, id)) return false;
   return true;
}

static std::string EmitKernel(GmlCtx *g, const char *KrnSrc, const char *PrcNam, int MshTyp,
                              const std::vector<Arg> &args, const char *ParSrc, bool lazy, bool stage,
                              bool pair = false, bool tile = false)
{
   std::string s;
   struct Local { std::string type, name, dims; };
   std::vector<Local> locals;
   auto local = [&](const std::string &t, const std::string &n, const std::string &d) {
      locals.push_back({t, n, d});
      return Fmt("   %s %s%s;\n", t.c_str(), n.c_str(), d.c_str());
   };
   s.reserve(16384);
   std::vector<int> staged, tstaged;
   const int sball = stage ? StagedArgs(args, &staged) : -1;
   const int stab = tile ? TileStagedArgs(g, args, MshTyp, KrnSrc, &tstaged) : -1;
   auto is_tstaged = [&](int i) { return std::find(tstaged.begin(), tstaged.end(), i) != tstaged.end(); };
   // the ball's own entries are read only when something needs them
   bool ball_ids = true;
   if(sball >= 0)
   {
      const Arg &B = args[sball];
      ball_ids = (B.FlgTab & GmlVoyeurs) || BodyUses(KrnSrc, B.nam);
      for(size_t i = 0; i < args.size(); i++)
         if(args[i].LnkIdx == sball && std::find(staged.begin(), staged.end(), (int)i) == staged.end())
            ball_ids = true;
   }
   s += "// GMlib MESHING AND GEOMETRICAL TOOLKIT\n";
   s += kPrelude;
   if(lazy)
      s += "#define GML_LAZY_UPLINKS 1\n"
           "template<class T, int NI> struct gml_uprow {\n"
           "   T (*tab)[NI]; int idx; bool ok; T nul;\n"
           "   __device__ __forceinline__ T operator[](int j) const { return ok ? tab[idx][j] : nul; }\n"
           "};\n"
           "template<class T, int NI, int W> struct gml_upn {\n"
           "   T (*tab)[NI]; const int *idx; int deg; T nul;\n"
           "   __device__ __forceinline__ gml_uprow<T, NI> operator[](int k) const { return gml_uprow<T, NI>{tab, idx[k], k < deg, nul}; }\n"
           "};\n";
   if(lazy && sball >= 0)
      // staged single-item data: each read from LDS where the body uses it
      s += "template<class T, int W> struct gml_stg {\n"
           "   const T *s; const unsigned short *l; int deg; T nul;\n"
           "   __device__ __forceinline__ T operator[](int k) const { return k < deg ? s[l[k]] : nul; }\n"
           "};\n";
   if(g->UsrTlk)
   {
      // OpenCL-C functions carry no qualifier: compiled for host and device
      // alike, so the kernel can call them
      s += "// USER'S TOOLKIT\n#pragma clang force_cuda_host_device begin\n";
      s += UserCode(g->UsrTlk);
      s += "\n#pragma clang force_cuda_host_device end\n";
   }
   s += "// USER'S ARGUMENTS STRUCTURE\n";
   if(ParSrc)
      s += UserCode(ParSrc);
   else
      s += "typedef struct { int GmlNoParameters; } GmlParSct;";
   s += "\n";

   // --- header
   s += "// KERNEL HEADER\n";
   s += Fmt("extern \"C\" __global__ void __launch_bounds__(256) %s(", PrcNam);
   for(auto &a : args)
      s += "\n   " + ArgDecl(a) + ",";
   if(sball >= 0)
      s += "\n   const unsigned int *gml_tofs,\n   const int *gml_tlist,\n   const unsigned short *gml_lidx,"
           "\n   const int gml_stage_n,";
   if(stab >= 0)
      s += "\n   const unsigned int *gml_dofs,\n   const int *gml_dlist,\n   const unsigned int *gml_dlid,"
           "\n   const int gml_dstage_n,\n   const int gml_drows,";
   s += "\n   GmlParSct *GmlPar,\n   const int2 count,\n   const int gml_remap)\n{\n";

   // --- locals (gmlib3.c:2670-2732)
   std::string decl, defs;
   for(size_t i = 0; i < args.size(); i++)
   {
      const Arg &a = args[i];
      if(a.FlgTab & GmlManual) continue;
      const Arg *L = a.LnkIdx >= 0 ? &args[a.LnkIdx] : nullptr;
      const Arg *C = a.CntIdx >= 0 ? &args[a.CntIdx] : nullptr;
      const char *T = OclTypStr[a.ItmTyp];
      const bool sd = std::find(staged.begin(), staged.end(), (int)i) != staged.end();
      bool lz = lazy && C && (a.NmbItm > 1 || sd);
      if(a.kind == AkBall)
      {
         if((int)i == sball) decl += local("unsigned short", "gml_l_" + a.nam, Fmt("[%d]", a.NmbItm * a.ItmLen));
         decl += local("int", "gml_" + a.nam, Fmt("[%d]", a.NmbItm * a.ItmLen));  // shifted entries, one per slot
         if(a.NmbItm > 1) decl += local("int16", a.nam, Fmt("[%d]", a.NmbItm));
         else             decl += local(OclTypStr[a.ItmTyp], a.nam, "");
      }
      else if(lz)
         ;  // declared after the reads as an accessor
      else if(L && L->MaxDeg > 1)
      {
         if(a.NmbItm > 1) decl += local(T, a.nam, Fmt("[%d][%d]", L->MaxDeg, a.NmbItm));
         else             decl += local(T, a.nam, Fmt("[%d]", L->MaxDeg));
      }
      else
      {
         if(a.NmbItm > 1) decl += local(T, a.nam, Fmt("[%d]", a.NmbItm));
         else             decl += local(T, a.nam, "");
      }
      if(C)
      {
         decl += local(OclTypStr[C->ItmTyp], a.nam + "Deg", "");
         decl += local(T, a.nam + "Nul", "");
         const std::string d = Fmt("   #define   %sDegMax %d\n", a.nam.c_str(), L->MaxDeg);
         decl += d;
         defs += d;
      }
      else if(L && L->LnkDir == 0)
         decl += local(T, a.nam + "Nul", "");
      if(a.FlgTab & GmlVoyeurs)
         decl += local("char     ", a.VoyNam, Fmt("[%d]", a.NmbItm * a.ItmLen));
   }

   // --- counters (gmlib3.c:2739-2748)
   const std::string idxnam = std::string(BalTypStr[MshTyp]) + "Idx";
   if(!pair && stab < 0)
   {
      s += "// KERNEL VARIABLES DEFINITION\n" + decl;
      s += "// KERNEL COUNTERS\n";
      s += "   int       cnt = gml_gid(gml_remap);\n";
      s += Fmt("   int       %s = cnt + count.s1;\n\n", idxnam.c_str());
   }
   // staged form: the tile's loads into LDS, emitted after the row's own
   // global reads so that both are in flight together
   std::string stage_s;
   if(sball >= 0)
   {
      std::swap(s, stage_s);
      s += "// STAGED UPLINK DATA: the work-group's tile, coalesced loads into LDS\n"
           "   extern __shared__ __align__(16) char gml_lds[];\n"
           "   size_t gml_lo = 0;\n";
      for(int i : staged)
      {
         const Arg &a = args[i];
         const char *T = OclTypStr[a.ItmTyp];
         s += Fmt("   %s *gml_s_%s = (%s *)(gml_lds + gml_lo);\n", T, a.nam.c_str(), T);
         s += Fmt("   gml_lo += ((size_t)gml_stage_n * sizeof(%s) + 15) & ~(size_t)15;\n", T);
      }
      // U entities per thread and pass, all U loads in flight before the
      // LDS stores; a slot past the end is clamped to the last entity, whose
      // value it loads and stores again (the same bytes, from cache).  With
      // 12 x 16 bytes a 256-row tile of a tet mesh's vertex balls (about
      // 2,300 distinct tets) takes one pass instead of two: one index ->
      // value round trip fewer (the drop-in smoothing gather at 256^3:
      // 721 -> 666 us); 10 to 12 measured alike, 14 and 16 slower (register
      // pressure), and loads predicated on the count instead of clamped
      // slower.  Wider items keep the same register budget.
      int bytes = 0;
      for(int i : staged) bytes += OclTypSiz[args[i].ItmTyp];
      const int U = bytes <= 16 ? 12 : (bytes <= 32 ? 6 : 3);
      s += Fmt("   {\n"
               "      const unsigned int gml_t = (unsigned int)(cnt - (int)threadIdx.x) / blockDim.x;\n"
               "      const unsigned int gml_t0 = gml_tofs[gml_t], gml_nt = gml_tofs[gml_t + 1] - gml_t0;\n"
               "      for(unsigned int gml_i = threadIdx.x; gml_i < gml_nt; gml_i += %d * blockDim.x)\n      {\n"
               "         int gml_x[%d];\n"
               "#pragma unroll\n"
               "         for(int gml_u = 0; gml_u < %d; gml_u++)\n"
               "            gml_x[gml_u] = gml_tlist[gml_t0 + min(gml_i + gml_u * blockDim.x, gml_nt - 1)];\n",
               U, U, U);
      for(int i : staged)
      {
         const char *n = args[i].nam.c_str();
         s += Fmt("         %s gml_v_%s[%d];\n#pragma unroll\n"
                  "         for(int gml_u = 0; gml_u < %d; gml_u++)\n"
                  "            gml_v_%s[gml_u] = %sTab[gml_x[gml_u]];\n",
                  OclTypStr[args[i].ItmTyp], n, U, U, n, n);
      }
      s += Fmt("#pragma unroll\n"
               "         for(int gml_u = 0; gml_u < %d; gml_u++)\n"
               "         {\n"
               "            const unsigned int gml_j = min(gml_i + gml_u * blockDim.x, gml_nt - 1);\n", U);
      for(int i : staged)
         s += Fmt("            gml_s_%s[gml_j] = gml_v_%s[gml_u];\n", args[i].nam.c_str(), args[i].nam.c_str());
      s += "         }\n      }\n      __syncthreads();\n   }\n\n";
      std::swap(s, stage_s);
   }
   if(!pair && sball < 0 && stab < 0) s += "   if(cnt >= count.s0)\n      return;\n\n";

   // --- reads (gmlib3.c:2755-2885), emitted into their own string
   std::string body_s, staged_reads;
   std::swap(s, body_s);
   s += "// KERNEL MEMORY READINGS\n";
   for(size_t i = 0; i < args.size(); i++)
   {
      const Arg &a = args[i];
      if(!(a.FlgTab & GmlReadMode) || (a.FlgTab & GmlManual)) continue;
      const Arg *L = a.LnkIdx >= 0 ? &args[a.LnkIdx] : nullptr;
      const Arg *C = a.CntIdx >= 0 ? &args[a.CntIdx] : nullptr;
      const char *n = a.nam.c_str();

      // element-tile form: the table row is not read; each slot of a staged
      // item comes from LDS through its byte of the row's tile-local indices
      if((int)i == stab) continue;
      if(stab >= 0 && is_tstaged((int)i))
      {
         const int siz = L->NmbItm * L->ItmLen;
         for(int k = 0; k < siz; k++)
            s += Fmt("   %s%s = gml_s_%s[(gml_lw[%d] >> %d) & 255u];\n", n, siz > 1 ? Fmt("[%d]", k).c_str() : "", n,
                     k >> 2, 8 * (k & 3));
         s += "\n";
         continue;
      }

      if(a.kind == AkBall)
      {
         // SELL-64 row of this entity; slots past the degree read as 0 as in
         // the reference's zero-padded rows.
         int w = a.NmbItm * a.ItmLen;
         const Arg *D = nullptr;
         for(auto &b : args) if(b.kind == AkDeg && b.SrcTyp == a.SrcTyp && b.DstTyp == a.DstTyp) D = &b;
         const bool st = (int)i == sball;
         s += Fmt("   {\n      const int gml_e = cnt + count.s1;\n"
                  "      const int *gml_row = %sTab + (size_t)%sOff[gml_e >> 6] * 64 + (gml_e & 63);\n"
                  "      const int gml_d = %sTab[gml_e];\n", n, n, D ? D->nam.c_str() : "");
         if(st)
            s += Fmt("      const unsigned short *gml_lrow = gml_lidx + (size_t)%sOff[gml_e >> 6] * 64 + (gml_e & 63);\n", n);
         s += Fmt("#pragma unroll\n      for(int k = 0; k < %d; k++)\n      {\n", w);
         if(st)
            s += Fmt("         gml_l_%s[k] = k < gml_d ? __builtin_nontemporal_load(&gml_lrow[k * 64]) : 0;\n", n);
         if(!st || ball_ids)
         {
            s += "         int v = k < gml_d ? __builtin_nontemporal_load(&gml_row[k * 64]) : 0;\n";
            if(a.FlgTab & GmlVoyeurs)
               s += Fmt("         %s[k] = (char)(v & 7);\n", a.VoyNam.c_str());
            s += Fmt("         gml_%s[k] = v >> 4;\n", n);
         }
         s += "      }\n   }\n";
         if(st && !ball_ids) { s += "\n"; continue; }
         for(int v = 0; v < a.NmbItm; v++)
         {
            if(a.NmbItm > 1) s += Fmt("   %s[%d] = (int16){", n, v);
            else             s += Fmt("   %s = (%s){", n, OclTypStr[a.ItmTyp]);
            for(int l = 0; l < a.ItmLen; l++)
               s += Fmt("%sgml_%s[%d]", l ? "," : "", n, v * a.ItmLen + l);
            s += "};\n";
         }
         s += "\n";
         continue;
      }
      if(a.kind == AkDeg)
      {
         s += Fmt("   %s = __builtin_nontemporal_load(&%sTab[ cnt + count.s1 ]);\n\n", n, n);
         continue;
      }

      if(C)
      {
         s += Fmt("   %sDeg = %s;\n", n, C->nam.c_str());
         s += Fmt("   %sNul = %s;\n", n, OclNulVec[a.ItmTyp]);
         if(lazy && a.NmbItm > 1)
         {
            s += Fmt("   gml_upn<%s, %d, %d> %s{%sTab, gml_%s, %sDeg, %sNul};\n\n",
                     OclTypStr[a.ItmTyp], a.NmbItm, L->MaxDeg, n, n, L->nam.c_str(), n, n);
            continue;
         }
      }
      else if(L && L->LnkDir == 0)
         s += Fmt("   %sNul = %s;\n", n, OclNulVec[a.ItmTyp]);

      int siz = (L && L->NmbItm * L->ItmLen > 1) ? L->NmbItm * L->ItmLen : 1;
      if(L && L->kind == AkBall) siz = L->MaxDeg;
      for(int j = 0; j < a.NmbItm; j++)
      {
         std::string td1 = a.NmbItm > 1 ? Fmt("[%d]", j) : "";
         for(int k = 0; k < siz; k++)
         {
            std::string td2 = siz > 1 ? Fmt("[%d]", k) : "";
            std::string lnk;
            if(!L)
               lnk = "cnt + count.s1";
            else if(L->kind == AkBall)
               lnk = Fmt("gml_%s[%d]", L->nam.c_str(), k);
            else
            {
               std::string l1 = L->NmbItm > 1 ? Fmt("[%d]", k / 16) : "";
               std::string l2 = L->ItmLen > 1 ? Fmt(".s%c", HexDig[k & 15]) : "";
               lnk = L->nam + l1 + l2;
            }
            const bool sd = std::find(staged.begin(), staged.end(), (int)i) != staged.end();
            if(C && sd && lazy)
            {
               if(k == 0)
                  staged_reads += Fmt("   gml_stg<%s, %d> %s{gml_s_%s, gml_l_%s, %sDeg, %sNul};\n", OclTypStr[a.ItmTyp],
                                      siz, n, n, L->nam.c_str(), n, n);
            }
            else if(C && sd)
               staged_reads += Fmt("   %s%s = (%sDeg >= %d) ? gml_s_%s[ gml_l_%s[%d] ] : %sNul;\n",
                                   n, td2.c_str(), n, k + 1, n, L->nam.c_str(), k, n);
            else if(C)
               s += Fmt("   %s%s%s = (%sDeg >= %d) ? %sTab[ %s ]%s : %sNul;\n",
                        n, td2.c_str(), td1.c_str(), n, k + 1, n, lnk.c_str(), td1.c_str(), n);
            else if(L && L->LnkDir == 0)
               s += Fmt("   %s%s%s = %s ? %sTab[ %s ]%s : %sNul;\n",
                        n, td2.c_str(), td1.c_str(), lnk.c_str(), n, lnk.c_str(), td1.c_str(), n);
            else if(!L)   // the entity's own line: read once, streamed past the caches
               s += Fmt("   %s%s%s = __builtin_nontemporal_load(&%sTab[ %s ]%s);\n", n, td2.c_str(), td1.c_str(), n,
                        lnk.c_str(), td1.c_str());
            else
               s += Fmt("   %s%s%s = %sTab[ %s ]%s;\n", n, td2.c_str(), td1.c_str(), n, lnk.c_str(), td1.c_str());
         }
      }
      s += "\n";
   }

   std::string reads;
   std::swap(s, reads);
   s = body_s;

   // --- body
   const std::string user = "// USER'S KERNEL CODE\n" + UserCode(KrnSrc) + "\n";

   // --- writes (gmlib3.c:2903-2938): only the loop entity's own line
   std::string writes_s;
   std::swap(s, writes_s);
   s += "\n// KERNEL MEMORY WRITINGS\n";
   for(auto &a : args)
   {
      if(!(a.FlgTab & GmlWriteMode) || (a.FlgTab & GmlManual)) continue;
      if(a.NmbItm == 1)
         s += Fmt("   __builtin_nontemporal_store(%s, &%sTab[ cnt + count.s1 ]);\n", a.nam.c_str(), a.nam.c_str());
      else
         for(int c = 0; c < a.NmbItm; c++)
            s += Fmt("   __builtin_nontemporal_store(%s[%d], &%sTab[ cnt + count.s1 ][%d]);\n", a.nam.c_str(), c,
                     a.nam.c_str(), c);
   }
   std::string writes;
   std::swap(s, writes);
   s = writes_s;
   if(!pair && sball >= 0)
   {
      // the row's global reads (a thread past the end reads the last row
      // and later returns: every thread reaches the barrier, and no branch
      // makes the compiler wait for these loads before the staging's), the
      // tile staging, then the reads of the staged data from LDS
      s += "   const int gml_cnt = cnt;\n   cnt = min(cnt, count.s0 - 1);\n" + reads + "   cnt = gml_cnt;\n\n" + stage_s +
           "   if(cnt >= count.s0)\n      return;\n\n// STAGED READINGS\n" + staged_reads + "\n" + user + writes + "}\n";
      return s;
   }
   if(stab >= 0)
   {
      // element-tile form: the work-group's tile of kDownTile rows, each
      // thread kDownTile / blockDim.x of them; the rows' index words first,
      // then the tile's distinct vertices into LDS, then the rows
      const int wq = (args[stab].NmbItm * args[stab].ItmLen + 3) / 4;
      s += Fmt("// ELEMENT TILE: gml_drows rows (at most %d), their distinct vertex data staged in LDS\n"
               "   const int gml_b = gml_blk(gml_remap);\n"
               "   const int gml_R = gml_drows / (int)blockDim.x;\n"
               "   unsigned int gml_lws[8][%d];\n"
               "#pragma unroll\n"
               "   for(int gml_q = 0; gml_q < 8; gml_q++)\n"
               "      if(gml_q < gml_R)\n      {\n"
               "         const int gml_r = min(gml_b * gml_drows + gml_q * (int)blockDim.x + (int)threadIdx.x, count.s0 - 1);\n"
               "#pragma unroll\n"
               "         for(int gml_j = 0; gml_j < %d; gml_j++)\n"
               "            gml_lws[gml_q][gml_j] = __builtin_nontemporal_load(&gml_dlid[(size_t)gml_r * %d + gml_j]);\n"
               "      }\n"
               "   extern __shared__ __align__(16) char gml_lds[];\n"
               "   size_t gml_lo = 0;\n", kDownTile, wq, wq, wq);
      for(int i : tstaged)
      {
         const char *T = OclTypStr[args[i].ItmTyp];
         s += Fmt("   %s *gml_s_%s = (%s *)(gml_lds + gml_lo);\n", T, args[i].nam.c_str(), T);
         s += Fmt("   gml_lo += ((size_t)gml_dstage_n * sizeof(%s) + 15) & ~(size_t)15;\n", T);
      }
      // two entities per thread and pass (a tile has at most 256), both
      // loads in flight before the LDS stores; a slot past the end is
      // clamped to the last entity
      s += "   {\n"
           "      const unsigned int gml_t0 = gml_dofs[gml_b], gml_nt = gml_dofs[gml_b + 1] - gml_t0;\n"
           "      for(unsigned int gml_i = threadIdx.x; gml_i < gml_nt; gml_i += 2 * blockDim.x)\n      {\n"
           "         int gml_x[2];\n"
           "#pragma unroll\n"
           "         for(int gml_u = 0; gml_u < 2; gml_u++)\n"
           "            gml_x[gml_u] = gml_dlist[gml_t0 + min(gml_i + gml_u * blockDim.x, gml_nt - 1)];\n";
      for(int i : tstaged)
      {
         const char *n = args[i].nam.c_str();
         s += Fmt("         %s gml_v_%s[2];\n#pragma unroll\n"
                  "         for(int gml_u = 0; gml_u < 2; gml_u++)\n"
                  "            gml_v_%s[gml_u] = %sTab[gml_x[gml_u]];\n", OclTypStr[args[i].ItmTyp], n, n, n);
      }
      s += "#pragma unroll\n"
           "         for(int gml_u = 0; gml_u < 2; gml_u++)\n"
           "         {\n"
           "            const unsigned int gml_j = min(gml_i + gml_u * blockDim.x, gml_nt - 1);\n";
      for(int i : tstaged)
         s += Fmt("            gml_s_%s[gml_j] = gml_v_%s[gml_u];\n", args[i].nam.c_str(), args[i].nam.c_str());
      s += "         }\n      }\n      __syncthreads();\n   }\n\n";
      s += "// ONE ROW: its locals, reads, the user's code and its writes\n"
           "   auto gml_row = [&](int cnt, const unsigned int *gml_lw) {\n";
      s += Fmt("   int       %s = cnt + count.s1;\n", idxnam.c_str());
      s += decl + reads + user + writes + "   };\n";
      s += "#pragma unroll\n"
           "   for(int gml_q = 0; gml_q < 8; gml_q++)\n"
           "      if(gml_q < gml_R)\n      {\n"
           "         const int gml_c = gml_b * gml_drows + gml_q * (int)blockDim.x + (int)threadIdx.x;\n"
           "         if(gml_c < count.s0)\n"
           "            gml_row(gml_c, gml_lws[gml_q]);\n"
           "      }\n}\n";
      return s;
   }
   if(!pair)
   {
      s += reads + user + writes + "}\n";
      return s;
   }

   // paired form: the row struct, the read / body / write lambdas, two rows
   std::string bind = "      int &cnt = gml_x.cnt;\n      int &" + idxnam + " = gml_x." + idxnam + ";\n";
   s += "// ROW VARIABLES (two rows per thread)\n" + defs + "   struct gml_row_t {\n";
   for(const Local &l : locals)
   {
      s += Fmt("      %s %s%s;\n", l.type.c_str(), l.name.c_str(), l.dims.c_str());
      if(l.dims.empty()) bind += Fmt("      %s &%s = gml_x.%s;\n", l.type.c_str(), l.name.c_str(), l.name.c_str());
      else bind += Fmt("      %s (&%s)%s = gml_x.%s;\n", l.type.c_str(), l.name.c_str(), l.dims.c_str(), l.name.c_str());
   }
   s += "      int cnt, " + idxnam + ";\n      bool gml_done;\n   };\n";
   s += "   auto gml_read = [&](gml_row_t &gml_x) {\n" + bind + reads + "   };\n";
   s += "   auto gml_body = [&](gml_row_t &gml_x) {\n" + bind + "      gml_x.gml_done = false;\n" + user +
        "      gml_x.gml_done = true;\n   };\n";
   s += "   auto gml_write = [&](gml_row_t &gml_x) {\n" + bind + writes + "   };\n";
   s += "// KERNEL COUNTERS\n"
        "   const int gml_c0 = gml_gid2(gml_remap);\n"
        "   if(gml_c0 >= count.s0)\n      return;\n"
        "   const bool gml_has1 = gml_c0 + (int)blockDim.x < count.s0;\n"
        "   gml_row_t gml_r0, gml_r1;\n"
        "   gml_r0.cnt = gml_c0;\n"
        "   gml_r1.cnt = gml_has1 ? gml_c0 + (int)blockDim.x : gml_c0;\n";
   s += "   gml_r0." + idxnam + " = gml_r0.cnt + count.s1;\n   gml_r1." + idxnam + " = gml_r1.cnt + count.s1;\n";
   // both bodies run (the second on a repeat of the first row at the end of
   // the range): a conditional second body would let the compiler sink the
   // second row's loads below the first body
   s += "   gml_read(gml_r0);\n   gml_read(gml_r1);\n"
        "   gml_body(gml_r0);\n   gml_body(gml_r1);\n"
        "   if(gml_r0.gml_done) gml_write(gml_r0);\n"
        "   if(gml_has1 && gml_r1.gml_done) gml_write(gml_r1);\n}\n";
   return s;
}

// Compile the emitted source into a new kernel slot.
static int NewKernel(GmlCtx *g, const std::string &src, const char *PrcNam, std::string *log)
{
   int idx = 0;
   for(int i = 1; i <= GmlMaxKrn; i++)
      if(!g->krn[i].use) { idx = i; break; }
   if(!idx)
   {
      fprintf(stderr, "gml: all %d kernel slots are in use\n", GmlMaxKrn);
      return 0;
   }
   std::string code;
   if(!RtcCompile(src, PrcNam, g->cflags, &code, log))
      return 0;
   Kernel *k = &g->krn[idx];
   *k = Kernel();
   SetShapes(k);
   if(g->host_only)
   {
      k->use = true;
      k->name = PrcNam;
      k->src = src;
      if(idx > g->NmbKrn) g->NmbKrn = idx;
      return idx;
   }
   if(hipModuleLoadData(&k->mod, code.data()) != hipSuccess
   || hipModuleGetFunction(&k->fn, k->mod, PrcNam) != hipSuccess)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: loading the code object of kernel %s failed\n", PrcNam);
      if(k->mod) (void)hipModuleUnload(k->mod);
      *k = Kernel();
      return 0;
   }
   k->use = true;
   k->name = PrcNam;
   k->src = src;
   if(idx > g->NmbKrn) g->NmbKrn = idx;
   return idx;
}

// ---------------------------------------------------------------------------
// Launch
// ---------------------------------------------------------------------------
// launch shapes the tuner tries: {group size, XCD remap, form}; the paired
// shapes for kernels with a paired form, the element-tile shapes for
// kernels with that form (dropped at the first launch if the table has no
// tile plan).  The first shape is launched again last, once code and caches
// are warm.
void SetShapes(Kernel *k)
{
   if(k->rowids)
   {
      // OpenCL's work-item ids name the row: no XCD remap, one row per item
      k->shapes = {{256, 0, 0}, {128, 0, 0}, {64, 0, 0}};
      k->remap = 0;
      return;
   }
   k->shapes = {{256, 1, 0}, {128, 1, 0}, {64, 1, 0}, {256, 0, 0}, {128, 0, 0}, {64, 0, 0}};
   if(k->fn2 || !k->src2.empty())
      for(std::array<int, 3> sh : {std::array<int, 3>{256, 1, 1}, {128, 1, 1}, {256, 0, 1}, {128, 0, 1}})
         k->shapes.push_back(sh);
   if(k->fn3 || !k->src3.empty())
      for(std::array<int, 3> sh : {std::array<int, 3>{128, 1, 2}, {256, 1, 2}, {64, 1, 2}, {128, 0, 2}, {256, 0, 2}})
         k->shapes.push_back(sh);
}
static int NmbShapes(const Kernel *k) { return (int)k->shapes.size() + 1; }
static const int *ShapeAt(const Kernel *k, int i) { return k->shapes[i < NmbShapes(k) - 1 ? i : 0].data(); }

void HarvestEvents(std::vector<LaunchRec> &ev, double *acc)
{
   size_t keep = 0;
   for(size_t i = 0; i < ev.size(); i++)
   {
      const hipError_t q = hipEventQuery(ev[i].end);
      if(q == hipSuccess)
      {
         float ms = 0.f;
         if(hipEventElapsedTime(&ms, ev[i].beg, ev[i].end) == hipSuccess)
            *acc += ms * 1e-3;
         FreeRec(ev[i]);
      }
      else if(q == hipErrorNotReady)
         ev[keep++] = ev[i];
      // any other answer: the pair is unusable (its launch is not timed)
      // and is neither kept nor returned to the pool
   }
   (void)hipGetLastError();
   ev.resize(keep);
}

// Timing events are pooled per device: creating and destroying two events
// for every launch cost several microseconds of host time per launch, more
// than a small kernel runs.  A pool keeps at most kEvPoolMax events.
static const int kEvPoolDev = 64;
static const size_t kEvPoolMax = 8192;
static std::mutex g_ev_mtx;
static std::vector<hipEvent_t> g_ev_pool[kEvPoolDev];

static bool TakeEvent(int dev, hipEvent_t *e)
{
   {
      std::lock_guard<std::mutex> lk(g_ev_mtx);
      if(!g_ev_pool[dev].empty()) { *e = g_ev_pool[dev].back(); g_ev_pool[dev].pop_back(); return true; }
   }
   return hipEventCreate(e) == hipSuccess;
}

static void GiveEvent(int dev, hipEvent_t e)
{
   if(!e) return;
   {
      std::lock_guard<std::mutex> lk(g_ev_mtx);
      if(dev >= 0 && dev < kEvPoolDev && g_ev_pool[dev].size() < kEvPoolMax) { g_ev_pool[dev].push_back(e); return; }
   }
   (void)hipEventDestroy(e);
}

bool NewRec(LaunchRec *r)
{
   int dev = 0;
   if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kEvPoolDev) { (void)hipGetLastError(); return false; }
   r->dev = dev;
   if(!TakeEvent(dev, &r->beg)) { (void)hipGetLastError(); return false; }
   if(!TakeEvent(dev, &r->end)) { (void)hipGetLastError(); GiveEvent(dev, r->beg); return false; }
   return true;
}

void FreeRec(const LaunchRec &r)
{
   GiveEvent(r.dev, r.beg);
   GiveEvent(r.dev, r.end);
}

int RunKernel(GmlCtx *g, Kernel *k)
{
   // argument block: one device pointer per slot (a ball slot also passes its
   // chunk offsets), then GmlPar, count and the remap flag
   std::vector<void *> ptrs;
   ptrs.reserve(k->args.size() * 2 + 1);
   for(int di : k->args)
   {
      if(di < 1 || di > GmlMaxDat || !g->dat[di].use || !g->dat[di].dev)
      {
         fprintf(stderr, "gml: invalid argument data %d of kernel %s\n", di, k->name.c_str());
         return -1;
      }
      ptrs.push_back(g->dat[di].dev);
      if(g->dat[di].kind == DkBal)
         ptrs.push_back(g->dat[g->dat[di].OffIdx].dev);
   }
   void *par = g->ParIdx && g->dat[g->ParIdx].dev ? g->dat[g->ParIdx].dev : g->dummy_par;
   long long n = k->NmbLin[0];
   if(n <= 0) return 1;
   int count[2] = {(int)k->NmbLin[0], (int)k->NmbLin[1]};

   // element-tile form: the table's tile plan, if it has one; without it
   // the form's shapes leave the tuner, and a kernel tuned to the form
   // falls back to one row per thread
   DownPlan *dp = nullptr;
   if(k->fn3)
   {
      dp = GetDownPlan(g, k->dtyp);
      if(dp && (dp->nrows != n || k->NmbLin[1] != 0)) dp = nullptr;
      if(!dp)
      {
         if(!k->tuned && k->tune_step == 0)
            k->shapes.erase(std::remove_if(k->shapes.begin(), k->shapes.end(),
                                           [](const std::array<int, 3> &x) { return x[2] == 2; }), k->shapes.end());
         if(k->tuned && k->form == 2) { k->form = 0; k->grp = 256; k->remap = 1; }
      }
   }
   int grp = k->grp, remap = k->remap, form = k->form, tuning = 0;
   if(!k->tuned)
   {
      const int *sh = ShapeAt(k, k->tune_step);
      grp = sh[0];
      remap = sh[1];
      form = sh[2] == 1 ? (k->fn2 ? 1 : 0) : (sh[2] == 2 && dp ? 2 : 0);
      tuning = 1;
   }
   const int pair = form == 1;
   // staged uplink data: the tile plan of this work-group size; the tuner
   // keeps the largest size whose LDS fits 64 KB and varies the remap only
   StagePlan *sp = nullptr;
   unsigned int lds = 0;
   int stage_n = 0;
   if(k->stage_src >= 0)
   {
      auto plan_for = [&](int gs) -> StagePlan * {
         StagePlan *&slot = g->up[k->stage_src][k->stage_dst].stage[gs == 64 ? 0 : (gs == 128 ? 1 : 2)];
         if(slot && slot->nrows != n) { FreeStagePlan(g, slot); slot = nullptr; }
         if(!slot) slot = NewStagePlan(g, k->stage_src, k->stage_dst, gs, n);
         return slot;
      };
      auto lds_of = [&](const StagePlan *p) {
         size_t b = 0;
         for(int z : k->stage_siz) b += ((size_t)p->max_nt * z + 15) & ~(size_t)15;
         return b;
      };
      if(!k->stage_grp)
         for(int gs : {256, 128, 64})
         {
            StagePlan *p = plan_for(gs);
            if(p && lds_of(p) <= 64 * 1024) { k->stage_grp = gs; break; }
         }
      if(!k->stage_grp)
      {
         fprintf(stderr, "gml: kernel %s: no tile of its staged uplink data fits the LDS\n", k->name.c_str());
         return -7;
      }
      if(tuning || !k->grp) grp = k->stage_grp;
      sp = plan_for(grp);
      if(!sp || lds_of(sp) > 64 * 1024)
      {
         fprintf(stderr, "gml: kernel %s: work-group size %d leaves its staged data too large for the LDS\n",
                 k->name.c_str(), grp);
         return -7;
      }
      lds = (unsigned int)lds_of(sp);
      stage_n = sp->max_nt;
   }
   int remap_arg = remap;
   std::vector<void *> kp;
   for(auto &p : ptrs) kp.push_back((void *)&p);
   void *stage_ptr[3] = {sp ? (void *)sp->tofs : nullptr, sp ? (void *)sp->tlist : nullptr,
                         sp ? (void *)sp->lidx : nullptr};
   if(sp)
   {
      for(auto &p : stage_ptr) kp.push_back((void *)&p);
      kp.push_back((void *)&stage_n);
   }
   void *tile_ptr[3] = {dp ? (void *)dp->tofs : nullptr, dp ? (void *)dp->tlist : nullptr, dp ? (void *)dp->lid : nullptr};
   int tile_n = dp ? dp->max_nv : 0, tile_rows = dp ? dp->rows : 0;
   if(form == 2)
   {
      // a work-group covers one tile: at most its rows of threads, at
      // most 8 rows per thread
      grp = std::min(grp, tile_rows);
      if(tile_rows / grp > 8) grp = tile_rows / 8;
      for(auto &p : tile_ptr) kp.push_back((void *)&p);
      kp.push_back((void *)&tile_n);
      kp.push_back((void *)&tile_rows);
      size_t b = 0;
      for(int z : k->dsiz) b += ((size_t)tile_n * z + 15) & ~(size_t)15;
      lds = (unsigned int)b;
   }
   kp.push_back((void *)&par);
   kp.push_back((void *)count);
   kp.push_back((void *)&remap_arg);

   LaunchRec rec, tr;
   bool have = NewRec(&rec);
   bool tun = tuning && NewRec(&tr);
   if(have) (void)hipEventRecord(rec.beg, g->stream);
   if(tun) (void)hipEventRecord(tr.beg, g->stream);
   const long long span = form == 2 ? tile_rows : (pair ? 2LL * grp : grp);
   unsigned int nb = (unsigned int)((n + span - 1) / span);
   hipFunction_t fn = form == 2 ? k->fn3 : (pair ? k->fn2 : k->fn);
   hipError_t e = hipModuleLaunchKernel(fn, nb, 1, 1, grp, 1, 1, lds, g->stream, kp.data(), nullptr);
   if(e != hipSuccess)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: launching kernel %s failed: %s\n", k->name.c_str(), hipGetErrorString(e));
      if(have) FreeRec(rec);
      if(tun) FreeRec(tr);
      return -6;
   }
   if(tun) (void)hipEventRecord(tr.end, g->stream);
   if(have)
   {
      (void)hipEventRecord(rec.end, g->stream);
      k->evts.push_back(rec);
      if(k->evts.size() > 256) HarvestEvents(k->evts, &k->run_time);
   }

   if(tuning)
   {
      if(tun) k->tune_ev.push_back(tr);
      k->tune_step++;
      if(k->tune_step == NmbShapes(k))
      {
         const int kNmbShapes = NmbShapes(k);
         // every shape launched once: keep the fastest.  The first shape is
         // launched again last, once code and caches are warm, and only that
         // second timing of it counts.
         float best = 1e30f;
         int bi = kNmbShapes - 1;
         if(!k->tune_ev.empty()) (void)hipEventSynchronize(k->tune_ev.back().end);
         for(int i = 1; i < (int)k->tune_ev.size() && i < kNmbShapes; i++)
         {
            float ms = 0.f;
            if(hipEventElapsedTime(&ms, k->tune_ev[i].beg, k->tune_ev[i].end) == hipSuccess && ms < best)
            { best = ms; bi = i; }
         }
         (void)hipGetLastError();
         for(auto &t : k->tune_ev) FreeRec(t);
         k->tune_ev.clear();
         const int *sh = ShapeAt(k, bi);
         k->grp = k->stage_src >= 0 ? k->stage_grp : sh[0];
         k->remap = sh[1];
         k->form = sh[2] == 1 ? (k->fn2 ? 1 : 0) : (sh[2] == 2 && k->fn3 ? 2 : 0);
         k->tuned = 1;
      }
   }

   // data written by the kernel: host mirrors are stale, versions new
   for(int di : k->wargs)
   {
      g->dat[di].host_valid = false;
      BumpVersion(&g->dat[di]);
   }
   return 1;
}

} // namespace gml

using namespace gml;

// ---------------------------------------------------------------------------
// GmlCompileKernel: classify arguments exactly as gmlib3.c:2238-2499, build
// missing links, emit and compile; add the high-degree twin when the uplink
// of the kernel has one.
// ---------------------------------------------------------------------------
static int CompileKernelTab(GmlCtx *g, const char *KrnSrc, const char *PrcNam, int MshTyp,
                            int NmbTyp, const int *IdxTab, const int *FlgIn, const int *LnkIn)
{
   if(!KrnSrc || !PrcNam || MshTyp < 0 || MshTyp >= GmlMaxEleTyp || NmbTyp < 0 || NmbTyp > GmlMaxDat)
      return 0;
   std::vector<int> FlgTab(FlgIn, FlgIn + NmbTyp), LnkTab(LnkIn, LnkIn + NmbTyp), CntTab(NmbTyp, 0);
   for(int i = 0; i < NmbTyp; i++)
      if(IdxTab[i] < 1 || IdxTab[i] > GmlMaxDat || !g->dat[IdxTab[i]].use)
      {
         fprintf(stderr, "gml: kernel %s: invalid data index %d\n", PrcNam, IdxTab[i]);
         return 0;
      }
   const char *ParSrc = g->ParIdx ? g->dat[g->ParIdx].src : nullptr;

   // 1. links between the kernel's entity kind and each argument's kind
   std::vector<int> UpSrc(NmbTyp, -1), UpDst(NmbTyp, -1);
   for(int i = 0; i < NmbTyp; i++)
   {
      Data *d = &g->dat[IdxTab[i]];
      if(LnkTab[i] || d->MshTyp == MshTyp) continue;
      int Dst = d->MshTyp;
      if(MshTypDim[MshTyp] > MshTypDim[Dst])
      {
         if(!g->LnkMat[MshTyp][Dst] && !BuildDownlink(g, MshTyp, Dst)) return 0;
         LnkTab[i] = g->LnkMat[MshTyp][Dst];
      }
      else if(MshTypDim[MshTyp] < MshTypDim[Dst])
      {
         if(!g->up[MshTyp][Dst].BalIdx && !BuildUplink(g, MshTyp, Dst)) return 0;
         g->up[MshTyp][Dst].pinned = true;   // the kernel holds its slots
         LnkTab[i] = g->up[MshTyp][Dst].BalIdx;
         CntTab[i] = g->up[MshTyp][Dst].DegIdx;
         UpSrc[i] = MshTyp; UpDst[i] = Dst;
      }
      else
      {
         fprintf(stderr, "gml: kernel %s: data %d of an entity kind of the same dimension needs an explicit link\n",
                 PrcNam, IdxTab[i]);
         return 0;
      }
   }

   // 2. context arguments (links first; an uplink is followed by its degree)
   std::vector<Arg> args;
   int HghSrc = -1, HghDst = -1;
   for(int i = 0; i < NmbTyp; i++)
   {
      if(!LnkTab[i]) continue;
      bool seen = false;
      for(auto &a : args) if(a.DatIdx == LnkTab[i]) seen = true;
      if(seen) continue;
      Data *d = &g->dat[IdxTab[i]];
      int Dst = d->MshTyp;
      Data *ld = &g->dat[LnkTab[i]];
      Arg a;
      a.MshTyp = Dst;
      a.DatIdx = LnkTab[i];
      a.FlgTab = GmlReadMode;
      a.nam = ld->nam;
      if(UpSrc[i] >= 0)
      {
         Uplink *u = &g->up[MshTyp][Dst];
         a.kind = AkBall;
         a.LnkDir = 1;
         a.MaxDeg = u->BasWid;
         int cnt, vec, typ;
         GetCntVec(u->BasWid, &cnt, &vec, &typ);
         a.NmbItm = cnt; a.ItmLen = vec; a.ItmTyp = typ;
         a.SrcTyp = MshTyp; a.DstTyp = Dst;
         if(u->HghWid)
         {
            if(HghSrc >= 0 && (HghSrc != MshTyp || HghDst != Dst))
            {
               fprintf(stderr, "gml: kernel %s mixes two uplinks with high-degree rows\n", PrcNam);
               return 0;
            }
            HghSrc = MshTyp; HghDst = Dst;
         }
         if(FlgTab[i] & GmlVoyeurs)
         {
            a.VoyNam = ld->VoyNam;
            a.FlgTab |= GmlVoyeurs;
            FlgTab[i] &= ~GmlVoyeurs;
         }
         args.push_back(a);
         Arg c;
         c.kind = AkDeg;
         c.MshTyp = Dst;
         c.DatIdx = u->DegIdx;
         c.FlgTab = GmlReadMode;
         c.ItmTyp = GmlInt;
         c.LnkDeg = c.MaxDeg = 1;
         c.nam = g->dat[u->DegIdx].nam;
         c.SrcTyp = MshTyp; c.DstTyp = Dst;
         args.push_back(c);
      }
      else
      {
         // downlink (element table or built link) or an explicit user link
         a.kind = (MshTypDim[MshTyp] == MshTypDim[Dst]) ? AkTrans : AkDown;
         a.LnkDir = (a.kind == AkTrans) ? 0 : -1;
         a.NmbItm = ld->NmbItm;
         a.ItmLen = ld->ItmLen;
         a.ItmTyp = ld->ItmTyp;
         a.LnkDeg = a.MaxDeg = ld->NmbItm * ld->ItmLen;
         args.push_back(a);
      }
   }

   // 3. data arguments, each optionally followed by its ref table
   for(int i = 0; i < NmbTyp; i++)
   {
      Data *d = &g->dat[IdxTab[i]];
      int LnkPos = -1, CptPos = -1;
      for(int j = 0; j < (int)args.size(); j++)
      {
         if(LnkTab[i] && args[j].DatIdx == LnkTab[i]) LnkPos = j;
         if(CntTab[i] && args[j].DatIdx == CntTab[i]) CptPos = j;
      }
      Arg a;
      a.MshTyp = d->MshTyp;
      a.DatIdx = IdxTab[i];
      a.LnkDir = 0;
      a.LnkIdx = LnkPos;
      a.CntIdx = CptPos;
      if(LnkPos >= 0) { a.MaxDeg = args[LnkPos].MaxDeg; a.LnkDeg = CptPos >= 0 ? -1 : args[LnkPos].LnkDeg; }
      bool RefFlg = (FlgTab[i] & GmlRefFlag) != 0;
      FlgTab[i] &= ~GmlRefFlag;
      a.NmbItm = d->NmbItm;
      a.ItmLen = d->ItmLen;
      a.ItmTyp = d->ItmTyp;
      a.FlgTab = FlgTab[i];
      a.nam = d->nam;
      args.push_back(a);
      if(!RefFlg || CptPos >= 0) continue;
      int RefIdx = g->RefIdx[d->MshTyp];
      if(!RefIdx) continue;
      Arg r;
      r.MshTyp = d->MshTyp;
      r.DatIdx = RefIdx;
      r.LnkIdx = LnkPos;
      r.CntIdx = -1;
      r.MaxDeg = LnkPos >= 0 ? args[LnkPos].MaxDeg : -1;
      r.NmbItm = 1; r.ItmLen = 0; r.ItmTyp = GmlInt;
      r.FlgTab = FlgTab[i] & ~GmlWriteMode;
      r.nam = g->dat[RefIdx].nam;
      args.push_back(r);
   }

   // 4. emit and compile: lazy uplink accessors first, eager arrays if the
   //    body needs them as lvalues
   bool any_lazy = false;
   for(auto &a : args) if(a.CntIdx >= 0 && a.NmbItm > 1) any_lazy = true;
   std::vector<int> staged;
   const int sball = g->stage_uplinks ? StagedArgs(args, &staged) : -1;
   std::string log, src;
   int KrnIdx = 0;
   for(int st = sball >= 0 ? 1 : 0; st >= 0 && !KrnIdx; st--)
   {
      if(any_lazy || st)
      {
         src = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, true, st);
         KrnIdx = NewKernel(g, src, PrcNam, &log);
      }
      if(!KrnIdx)
      {
         log.clear();
         src = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, false, st);
         KrnIdx = NewKernel(g, src, PrcNam, &log);
      }
      if(KrnIdx && st)
      {
         Kernel *k = &g->krn[KrnIdx];
         k->stage_src = args[sball].SrcTyp;
         k->stage_dst = args[sball].DstTyp;
         for(int i : staged) k->stage_siz.push_back(OclTypSiz[args[i].ItmTyp]);
      }
   }
   if(!KrnIdx)
   {
      printf("Compiling the kernel %s failed\n%s\n", PrcNam, log.c_str());
      fflush(stdout);
      if(g->DbgFlg) puts(src.c_str());
      return 0;
   }
   DbgPrint(g, "Generated source for kernel=%s, index=%2d\n%s\n", PrcNam, KrnIdx, g->DbgFlg ? src.c_str() : "");

   Kernel *k = &g->krn[KrnIdx];
   // the paired form (two rows per thread) of a kernel without uplinks; the
   // launch tuner chooses between the two forms
   if(HghSrc < 0 && k->stage_src < 0 && PairEligible(args, KrnSrc, false, false)
   && k->src.find("gml_upn<") == std::string::npos)
   {
      std::string psrc = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, false, false, true), code, plog;
      if(RtcCompile(psrc, PrcNam, g->cflags, &code, &plog))
      {
         k->src2 = psrc;
         if(!g->host_only
         && (hipModuleLoadData(&k->mod2, code.data()) != hipSuccess
             || hipModuleGetFunction(&k->fn2, k->mod2, PrcNam) != hipSuccess))
         {
            (void)hipGetLastError();
            if(k->mod2) (void)hipModuleUnload(k->mod2);
            k->mod2 = nullptr;
            k->fn2 = nullptr;
            k->src2.clear();
         }
      }
      else
         DbgPrint(g, "Kernel %s: the paired form does not compile, the one-row form is kept\n%s\n", PrcNam,
                  plog.c_str());
   }
   // the element-tile form of an element loop reading vertex data through
   // its own table; the launch tuner compares it with the other forms
   std::vector<int> tstaged;
   if(HghSrc < 0 && k->stage_src < 0 && g->stage_uplinks && TileStagedArgs(g, args, MshTyp, KrnSrc, &tstaged) >= 0)
   {
      std::string tsrc = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, false, false, false, true), code, tlog;
      if(RtcCompile(tsrc, PrcNam, g->cflags, &code, &tlog))
      {
         k->src3 = tsrc;
         k->dtyp = MshTyp;
         for(int i : tstaged) k->dsiz.push_back(OclTypSiz[args[i].ItmTyp]);
         if(!g->host_only
         && (hipModuleLoadData(&k->mod3, code.data()) != hipSuccess
             || hipModuleGetFunction(&k->fn3, k->mod3, PrcNam) != hipSuccess))
         {
            (void)hipGetLastError();
            if(k->mod3) (void)hipModuleUnload(k->mod3);
            k->mod3 = nullptr;
            k->fn3 = nullptr;
            k->src3.clear();
            k->dtyp = -1;
            k->dsiz.clear();
         }
      }
      else
         DbgPrint(g, "Kernel %s: the element-tile form does not compile, the other forms are kept\n%s\n", PrcNam,
                  tlog.c_str());
   }
   k->rowids = UsesRowIds(KrnSrc);
   SetShapes(k);
   for(auto &a : args)
   {
      k->args.push_back(a.DatIdx);
      if(a.FlgTab & (GmlWriteMode | GmlManual)) k->wargs.push_back(a.DatIdx);
   }
   k->NmbLin[1] = 0;
   if(HghSrc < 0)
      k->NmbLin[0] = g->NmbEle[MshTyp];
   else
      k->NmbLin[0] = g->up[HghSrc][HghDst].MaxPos;
   if(HghSrc < 0) return KrnIdx;

   // 5. high-degree twin: same arguments, wider ball, rows MaxPos..N-1
   Uplink *u = &g->up[HghSrc][HghDst];
   for(auto &a : args)
      if(a.kind == AkBall && a.SrcTyp == HghSrc && a.DstTyp == HghDst)
      {
         int cnt, vec, typ;
         GetCntVec(u->HghWid, &cnt, &vec, &typ);
         a.MaxDeg = u->HghWid; a.NmbItm = cnt; a.ItmLen = vec; a.ItmTyp = typ;
      }
   for(auto &a : args)
      if(a.LnkIdx >= 0 && args[a.LnkIdx].kind == AkBall)
         a.MaxDeg = args[a.LnkIdx].MaxDeg;
   int HghIdx = 0;
   if(any_lazy)
   {
      src = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, true, false);
      HghIdx = NewKernel(g, src, PrcNam, &log);
   }
   if(!HghIdx)
   {
      src = EmitKernel(g, KrnSrc, PrcNam, MshTyp, args, ParSrc, false, false);
      HghIdx = NewKernel(g, src, PrcNam, &log);
   }
   if(!HghIdx)
   {
      printf("Compiling the high-degree kernel %s failed\n%s\n", PrcNam, log.c_str());
      return 0;
   }
   Kernel *h = &g->krn[HghIdx];
   h->rowids = g->krn[KrnIdx].rowids;
   SetShapes(h);
   h->args = g->krn[KrnIdx].args;
   h->wargs = g->krn[KrnIdx].wargs;
   h->NmbLin[0] = g->NmbEle[MshTyp] - u->MaxPos;
   h->NmbLin[1] = u->MaxPos;
   g->krn[KrnIdx].HghIdx = HghIdx;
   return KrnIdx;
}

extern "C" {

int GmlCompileKernel(size_t GmlIdx, char *KrnSrc, char *PrcNam, int MshTyp, int NmbTyp, ...)
{
   GML_CTX(g, GmlIdx);
   if(NmbTyp < 0 || NmbTyp > GmlMaxDat) return 0;
   std::vector<int> idx(NmbTyp), flg(NmbTyp), lnk(NmbTyp);
   va_list ap;
   va_start(ap, NmbTyp);
   for(int i = 0; i < NmbTyp; i++)
   {
      idx[i] = va_arg(ap, int);
      flg[i] = va_arg(ap, int);
      lnk[i] = va_arg(ap, int);
   }
   va_end(ap);
   return CompileKernelTab(g, KrnSrc, PrcNam, MshTyp, NmbTyp, idx.data(), flg.data(), lnk.data());
}

int GmlxCompileKernelv(size_t GmlIdx, const char *KrnSrc, const char *PrcNam, int MshTyp,
                       int NmbTyp, const int *tri)
{
   GML_CTX(g, GmlIdx);
   if(NmbTyp < 0 || NmbTyp > GmlMaxDat || (NmbTyp && !tri)) return 0;
   std::vector<int> idx(NmbTyp), flg(NmbTyp), lnk(NmbTyp);
   for(int i = 0; i < NmbTyp; i++)
   {
      idx[i] = tri[3 * i];
      flg[i] = tri[3 * i + 1];
      lnk[i] = tri[3 * i + 2];
   }
   return CompileKernelTab(g, KrnSrc, PrcNam, MshTyp, NmbTyp, idx.data(), flg.data(), lnk.data());
}

int GmlLaunchKernel(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use || !g->krn[idx].fn)
      return -1;
   int res = RunKernel(g, &g->krn[idx]);
   if(res != 1) return res;
   if(g->krn[idx].HghIdx)
      res = RunKernel(g, &g->krn[g->krn[idx].HghIdx]);
   return res;
}

// Sum of the device durations of every launch of a kernel (and of its
// high-degree twin), in seconds.
double GmlGetKernelRunTime(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use || g->host_only)
      return -1.;
   HIPCHK(hipStreamSynchronize(g->stream));
   Kernel *k = &g->krn[idx];
   HarvestEvents(k->evts, &k->run_time);
   double t = k->run_time;
   if(k->HghIdx)
   {
      Kernel *h = &g->krn[k->HghIdx];
      HarvestEvents(h->evts, &h->run_time);
      t += h->run_time;
   }
   return t;
}

int GmlxSetStaging(size_t GmlIdx, int on)
{
   GML_CTX(g, GmlIdx);
   const int old = g->stage_uplinks;
   g->stage_uplinks = on ? 1 : 0;
   return old;
}

const char *GmlxGetKernelSource(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use) return nullptr;
   return g->krn[idx].src.c_str();
}

int GmlxGetKernelRows(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use) return -1;
   const Kernel *k = &g->krn[idx];
   if(k->src2.empty() && k->src3.empty()) return 1;
   if(!k->tuned) return 0;
   return k->form == 2 ? kDownTile / k->grp : (k->form == 1 ? 2 : 1);
}

int GmlxGetKernelForm(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use) return -1;
   const Kernel *k = &g->krn[idx];
   if(k->src2.empty() && k->src3.empty()) return 1;
   if(!k->tuned) return 0;
   return k->form + 1;
}

int GmlxSetKernelShape(size_t GmlIdx, int idx, int GrpSiz, int Remap, int Rows)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use) return 0;
   Kernel *k = &g->krn[idx];
   if(GrpSiz < 64 || GrpSiz > 256 || (GrpSiz & 63) || (Rows != 1 && Rows != 2 && Rows != -1)) return 0;
   if(Rows == 2 && !k->fn2 && (g->host_only ? k->src2.empty() : true)) return 0;
   if(Rows == -1 && !k->fn3 && (g->host_only ? k->src3.empty() : true)) return 0;
   if(k->stage_src >= 0 && GrpSiz != 64 && GrpSiz != 128 && GrpSiz != 256) return 0;
   if(Remap && k->rowids) return 0;          // work-item ids must name the row
   for(int i : {idx, k->HghIdx})
   {
      if(!i) continue;
      Kernel *t = &g->krn[i];
      for(auto &e : t->tune_ev) FreeRec(e);
      t->tune_ev.clear();
      t->grp = GrpSiz;
      if(t->stage_src >= 0) t->stage_grp = GrpSiz;
      t->remap = Remap ? 1 : 0;
      t->form = Rows == 2 && (t->fn2 || !t->src2.empty()) ? 1 : (Rows == -1 && (t->fn3 || !t->src3.empty()) ? 2 : 0);
      t->tuned = 1;
   }
   return 1;
}

const char *GmlxGetKernelPairSource(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use || g->krn[idx].src2.empty()) return nullptr;
   return g->krn[idx].src2.c_str();
}

const char *GmlxGetKernelTileSource(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use || g->krn[idx].src3.empty()) return nullptr;
   return g->krn[idx].src3.c_str();
}

int GmlxGetKernelShape(size_t GmlIdx, int idx, int *GrpSiz, int *Remap)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxKrn || !g->krn[idx].use) return 0;
   Kernel *k = &g->krn[idx];
   *GrpSiz = k->tuned ? k->grp : 0;
   *Remap = k->tuned ? k->remap : 0;
   return 1;
}


} // extern "C"

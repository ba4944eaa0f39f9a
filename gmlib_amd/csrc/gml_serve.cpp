// Serving loop for mesh jobs (the request side of the engine).
//
// A server owns a cache of device-resident meshes, each in its own GMlib
// handle (own HIP stream), and a pool of slots (worker threads).  Jobs name
// a seeded cube mesh (cells per side, seed) and a kind of work:
//   quality:  tet mean ratio + ball minimum + min / mean   (GmlxTetQuality,
//             GmlxBallMin, GmlReduceVector)
//   smooth:   nsweep Jacobi barycentre sweeps from the mesh's original
//             coordinates, then min / mean quality          (GmlxSmoothSweeps)
//   cg:       block-Jacobi CG with FP64 refinement on the mesh's edge
//             Laplacian (GmlxEdgeLaplacian3), hashed N(0,1) right-hand side
//                                                           (GmlxSolveCG)
// Admission is continuous: whenever a slot is free, the oldest pending job
// that can run is admitted -- its mesh resident and idle, or loadable within
// the HBM budget after evicting least-recently-used idle meshes -- so jobs
// arriving while others run start as soon as a slot and memory free up, and
// their kernels overlap on the device (one stream per resident mesh).  A
// mesh serves one job at a time; a smoothing job restores the original
// coordinates first (and so does any job after one), so every result is
// independent of what else ran: the same alone and in the mix.
//
// Meshes are either seeded cube meshes (cells per side, seed) or meshes the
// caller registered with GmlxServeAddMesh.  The server keeps the host copy
// of every mesh it has seen (a cube mesh is generated on its first job) and,
// once built, its CG matrix and right-hand side, so a mesh evicted from the
// device is loaded again by an upload alone.
//
// Resident = 0 gives the plain per-job path for comparison: every job loads
// its mesh into a fresh handle (GmlInit, upload, links and plans built on
// the device), runs, and frees it, one job at a time.
//
// Lifetime: a job's record lives until every caller blocked in
// GmlxServeWait on it has returned (the first to return retires its id);
// GmlxServeStop lets running jobs end, finishes pending ones with -8 and
// waits for blocked waiters to leave before freeing the server.

#include "gml_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>

namespace gml {

// host copy of a mesh (registered by the caller, or a generated cube mesh),
// with its CG system once a CG job has built it
struct UserMesh {
   int nver = 0, ntet = 0;
   unsigned long long seed = 0;    // seed of the CG matrix weights and right-hand side
   std::vector<float> crd;
   std::vector<int> ref, tet;
   std::mutex sys_mtx;
   long long nnz = 0;             // > 0 once the CG system is built
   std::vector<int> rowptr, col;
   std::vector<float> val, rhs;
};

// n > 0: cube mesh (n, seed); n == 0: user mesh number seed
struct Resident {
   int n = 0;
   unsigned long long seed = 0;
   size_t h = 0;
   int nver = 0, ntet = 0;
   int VerIdx = 0, TetIdx = 0, qal = 0, vqal = 0, bary = 0, crd0 = 0;
   int mat = 0, rhs = 0, sol = 0;
   std::shared_ptr<UserMesh> mesh;   // host copy
   size_t bytes = 0;              // device bytes in use (or reserved before the job ends)
   bool busy = false, dirty = false;
   long long last = 0;
};

struct ServeJob {
   long long id = 0;
   int type = 0, n = 0;
   unsigned long long seed = 0;
   double param = 0.;
   double t_submit = 0., t_admit = 0., t_done = 0.;
   double res[4] = {0., 0., 0., 0.};
   int rc = 0;
   bool loaded = false, done = false;
   int waiters = 0;               // callers blocked in GmlxServeWait on this job
   bool retired = false;          // id removed from the job map
};

struct Server {
   int dev = 0, nslots = 1;
   bool resident = true;
   size_t budget = 0, used = 0, peak = 0;
   std::mutex mtx;
   std::condition_variable cv_work, cv_done;
   std::deque<ServeJob *> pending;
   std::map<long long, ServeJob *> jobs;
   std::vector<Resident *> cache;
   std::map<unsigned long long, std::shared_ptr<UserMesh>> meshes;           // registered
   std::map<std::pair<int, unsigned long long>, std::shared_ptr<UserMesh>> cubes;   // generated
   unsigned long long next_mesh = 1;
   int nwaiters = 0;
   std::vector<std::thread> workers;
   bool stop = false;
   long long next_id = 1, clock = 0, loads = 0, evictions = 0, ndone = 0;
   int running = 0, max_running = 0;
   std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
   double Now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// vertex and tet counts of a job's mesh (lock held)
static void MeshSize(const Server *s, int n, unsigned long long seed, double *V, double *T)
{
   if(n > 0) { *V = std::pow(n + 1.0, 3); *T = 6.0 * std::pow((double)n, 3); return; }
   auto it = s->meshes.find(seed);
   *V = it == s->meshes.end() ? 0. : it->second->nver;
   *T = it == s->meshes.end() ? 0. : it->second->ntet;
}

// device bytes a job's mesh may need (upper estimate, before loading)
static size_t Estimate(const Server *s, int n, unsigned long long seed, int type)
{
   double V, T;
   MeshSize(s, n, seed, &V, &T);
   double b = 44 * T + 48 * V + (1 << 20);
   if(type == GmlxJobSmooth) b += 30 * T + 16 * V;
   if(type == GmlxJobCG) b += 900 * V;
   return (size_t)b;
}

static size_t DeviceBytes(const Resident *r) { return r->h ? GmlGetMemoryUsage(r->h) + (2 << 20) : 0; }

static void Unload(Resident *r)
{
   if(r->h) GmlStop(r->h);
   r->h = 0;
}

// the host copy of a cube mesh, generated on first use (lock not held)
static std::shared_ptr<UserMesh> CubeHost(Server *s, int n, unsigned long long seed)
{
   {
      std::lock_guard<std::mutex> lk(s->mtx);
      auto it = s->cubes.find({n, seed});
      if(it != s->cubes.end()) return it->second;
   }
   const long long V = (long long)(n + 1) * (n + 1) * (n + 1), T = 6LL * n * n * n;
   if(V >= (1LL << 31) || T >= (1LL << 31)) return nullptr;
   auto m = std::make_shared<UserMesh>();
   m->nver = (int)V;
   m->ntet = (int)T;
   m->seed = seed;
   m->crd.resize((size_t)V * 4);
   m->ref.resize((size_t)V);
   m->tet.resize((size_t)T * 4);
   if(!GmlxCubeMesh(n, 0.2, seed, 1, m->crd.data(), m->ref.data(), m->tet.data())) return nullptr;
   std::lock_guard<std::mutex> lk(s->mtx);
   auto ins = s->cubes.emplace(std::make_pair(n, seed), m);
   return ins.first->second;
}

// load the host mesh of r into a fresh handle on the server's device; false
// on failure
static bool Load(const Server *s, Resident *r)
{
   UserMesh *m = r->mesh.get();
   if(!m) return false;
   const long long V = m->nver, T = m->ntet;
   if(V < 1 || T < 1) return false;
   std::vector<int> tref((size_t)T, 0);
   r->nver = (int)V;
   r->ntet = (int)T;
   r->h = GmlInit(s->dev);
   if(!r->h) return false;
   const size_t h = r->h;
   float *crd = m->crd.data();
   int *ref = m->ref.data(), *tet = m->tet.data();
   r->VerIdx = GmlNewMeshData(h, GmlVertices, (int)V);
   r->TetIdx = GmlNewMeshData(h, GmlTetrahedra, (int)T);
   if(!r->VerIdx || !r->TetIdx
   || !GmlSetDataBlock(h, GmlVertices, 0, (int)V - 1, crd, crd + (V - 1) * 4, ref, ref + V - 1)
   || !GmlSetDataBlock(h, GmlTetrahedra, 0, (int)T - 1, tet, tet + (T - 1) * 4, tref.data(), tref.data() + T - 1))
      return false;
   r->qal = GmlNewSolutionData(h, GmlTetrahedra, 1, GmlFlt, (char *)"qal");
   r->vqal = GmlNewSolutionData(h, GmlVertices, 1, GmlFlt, (char *)"vqal");
   r->crd0 = GmlNewSolutionData(h, GmlVertices, 1, GmlFlt4, (char *)"crd0");
   if(!r->qal || !r->vqal || !r->crd0) return false;
   // pristine copy of the coordinates, restored before a job that needs it
   size_t nb = 0, nb0 = 0;
   void *src = GmlxGetDevicePointer(h, r->VerIdx, &nb), *dst = GmlxGetDevicePointer(h, r->crd0, &nb0);
   if(!src || !dst || nb0 < (size_t)V * 16) return false;
   if(hipMemcpyAsync(dst, src, (size_t)V * 16, hipMemcpyDeviceToDevice, (hipStream_t)GmlxGetStream(h)) != hipSuccess)
      return false;
   return GmlxSync(h) == 1;
}

static bool Restore(Resident *r)
{
   size_t nb = 0, nb0 = 0;
   void *dst = GmlxGetDevicePointer(r->h, r->VerIdx, &nb), *src = GmlxGetDevicePointer(r->h, r->crd0, &nb0);
   if(!src || !dst) return false;
   if(hipMemcpyAsync(dst, src, (size_t)r->nver * 16, hipMemcpyDeviceToDevice, (hipStream_t)GmlxGetStream(r->h)) != hipSuccess)
      return false;
   GmlxInvalidateHost(r->h, r->VerIdx);
   r->dirty = false;
   return true;
}

static int QualityStats(Resident *r, double *res)
{
   double mn = 0., sm = 0.;
   if(GmlxTetQuality(r->h, r->qal) != 1 || GmlxBallMin(r->h, r->qal, r->vqal) != 1) return -6;
   if(GmlReduceVector(r->h, r->qal, GmlMin, &mn) != 1 || GmlReduceVector(r->h, r->qal, GmlSum, &sm) != 1) return -6;
   res[0] = mn;
   res[1] = sm / (double)r->ntet;
   return 1;
}

// hashed N(0,1) entries (Box-Muller on two splitmix draws)
static double HashNormal(unsigned long long seed, long long i)
{
   auto mix = [](unsigned long long z) {
      z += 0x9e3779b97f4a7c15ULL;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      return z ^ (z >> 31);
   };
   const unsigned long long k = mix(seed ^ mix((unsigned long long)i));
   const double u1 = ((double)(mix(k + 1) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
   const double u2 = (double)(mix(k + 2) >> 11) * (1.0 / 9007199254740992.0);
   return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

static int RunJob(Resident *r, ServeJob *j)
{
   if(j->type == GmlxJobQuality)
   {
      if(r->dirty && !Restore(r)) return -6;
      return QualityStats(r, j->res);
   }
   if(j->type == GmlxJobSmooth)
   {
      if(!Restore(r)) return -6;
      if(!r->bary && !(r->bary = GmlNewSolutionData(r->h, GmlTetrahedra, 1, GmlFlt4, (char *)"bary"))) return -5;
      r->dirty = true;
      const int nsweep = j->param > 0 ? (int)j->param : 20;
      if(GmlxSmoothSweeps(r->h, r->bary, nsweep, 0.2) != 1) return -6;
      return QualityStats(r, j->res);
   }
   if(j->type == GmlxJobCG)
   {
      if(!r->mat)
      {
         UserMesh *m = r->mesh.get();
         const int V = m->nver;
         {
            // the host system is built once per mesh and kept with it
            std::lock_guard<std::mutex> lk(m->sys_mtx);
            if(m->nnz <= 0)
            {
               m->rowptr.resize((size_t)V + 1);
               const long long nnz = GmlxEdgeLaplacian3(V, m->ntet, m->tet.data(), m->seed, 0.01, m->rowptr.data(),
                                                        nullptr, nullptr, 0);
               if(nnz <= 0) return -5;
               m->col.resize((size_t)nnz);
               m->val.resize((size_t)nnz * 9);
               if(GmlxEdgeLaplacian3(V, m->ntet, m->tet.data(), m->seed, 0.01, m->rowptr.data(), m->col.data(),
                                     m->val.data(), nnz) != nnz)
                  return -5;
               m->rhs.resize((size_t)V * 3);
               for(long long i = 0; i < (long long)V * 3; i++) m->rhs[i] = (float)HashNormal(m->seed, i);
               m->nnz = nnz;
            }
         }
         r->mat = GmlNewMatrix(r->h, V, (int)m->nnz, 3, m->val.data(), m->col.data(), m->rowptr.data(), GmlFlt);
         r->rhs = GmlNewVector(r->h, V, 3, m->rhs.data(), GmlFlt);
         r->sol = GmlNewVector(r->h, V, 3, nullptr, GmlDbl);
         if(!r->mat || !r->rhs || !r->sol) return -5;
      }
      const double tol = j->param > 0 ? j->param : 1e-8;
      double rel = 0.;
      int it = 0;
      const int rc = GmlxSolveCG(r->h, r->mat, r->rhs, r->sol, tol, 20000, 1, &rel, &it);
      j->res[0] = rel;
      j->res[1] = it;
      return rc;
   }
   return -1;
}

// evict least-recently-used idle meshes (never keep) until need more bytes
// fit the budget (lock held)
static bool MakeRoom(Server *s, size_t need, const Resident *keep)
{
   while(s->used + need > s->budget)
   {
      Resident *v = nullptr;
      for(Resident *r : s->cache)
         if(!r->busy && r != keep && (!v || r->last < v->last)) v = r;
      if(!v) return false;
      s->used -= v->bytes;
      s->cache.erase(std::find(s->cache.begin(), s->cache.end(), v));
      Unload(v);
      delete v;
      s->evictions++;
   }
   return true;
}

static void Finish(Server *s, ServeJob *j, int rc)
{
   j->rc = rc;
   j->done = true;
   j->t_done = s->Now();
   s->ndone++;
   s->cv_done.notify_all();
}

// the oldest pending job that can run now, its mesh reserved (lock held)
static ServeJob *Admit(Server *s, Resident **out)
{
   for(auto it = s->pending.begin(); it != s->pending.end();)
   {
      ServeJob *j = *it;
      const size_t est = Estimate(s, j->n, j->seed, j->type);
      if(est > s->budget)
      {
         it = s->pending.erase(it);
         j->t_admit = s->Now();
         Finish(s, j, -7);
         continue;
      }
      if(!s->resident && s->running > 0) return nullptr;
      Resident *r = nullptr;
      if(s->resident)
         for(Resident *c : s->cache)
            if(c->n == j->n && c->seed == j->seed) { r = c; break; }
      if(r && r->busy) { ++it; continue; }
      const size_t need = r ? (est > r->bytes ? est - r->bytes : 0) : est;
      if(need && !MakeRoom(s, need, r)) { ++it; continue; }
      if(!r)
      {
         r = new Resident();
         r->n = j->n;
         r->seed = j->seed;
         if(j->n == 0)
         {
            auto mi = s->meshes.find(j->seed);
            if(mi != s->meshes.end()) r->mesh = mi->second;
         }
         s->cache.push_back(r);
         j->loaded = true;
         s->loads++;
      }
      // the reservation joins the mesh's bytes until its real size is known
      s->used += need;
      r->bytes += need;
      s->peak = std::max(s->peak, s->used);
      r->busy = true;
      r->last = ++s->clock;
      s->pending.erase(it);
      *out = r;
      return j;
   }
   return nullptr;
}

static void Worker(Server *s)
{
   (void)hipSetDevice(s->dev);
   for(;;)
   {
      std::unique_lock<std::mutex> lk(s->mtx);
      ServeJob *j = nullptr;
      Resident *r = nullptr;
      s->cv_work.wait(lk, [&]() { return s->stop || (j = Admit(s, &r)) != nullptr; });
      if(!j) return;
      s->running++;
      s->max_running = std::max(s->max_running, s->running);
      j->t_admit = s->Now();
      lk.unlock();

      int rc = 1;
      if(!r->mesh && r->n > 0) r->mesh = CubeHost(s, r->n, r->seed);
      if(!r->h && !Load(s, r)) rc = -5;
      if(rc == 1) rc = RunJob(r, j);

      lk.lock();
      const size_t real = DeviceBytes(r);
      if(rc == -5 || !s->resident)
      {
         // failed load, or the plain path: the mesh goes
         s->used -= r->bytes;
         s->cache.erase(std::find(s->cache.begin(), s->cache.end(), r));
         Unload(r);
         delete r;
      }
      else
      {
         s->used = s->used - r->bytes + real;
         r->bytes = real;
         r->busy = false;
         r->last = ++s->clock;
         s->peak = std::max(s->peak, s->used);
         MakeRoom(s, 0, nullptr);
      }
      s->running--;
      Finish(s, j, rc);
      s->cv_work.notify_all();
   }
}

} // namespace gml

using namespace gml;

extern "C" {

size_t GmlxServeStart(int dev, double BudgetGB, int NmbSlot, int Resident)
{
   if(BudgetGB <= 0. || NmbSlot < 1 || NmbSlot > 64) return 0;
   Server *s = new Server();
   s->dev = dev;
   s->budget = (size_t)(BudgetGB * 1e9);
   s->resident = Resident != 0;
   s->nslots = s->resident ? NmbSlot : 1;
   for(int i = 0; i < s->nslots; i++) s->workers.emplace_back(Worker, s);
   return (size_t)s;
}

static long long Enqueue(Server *s, int type, int n, unsigned long long seed, double param)
{
   ServeJob *j = new ServeJob();
   j->type = type;
   j->n = n;
   j->seed = seed;
   j->param = param;
   std::lock_guard<std::mutex> lk(s->mtx);
   j->id = s->next_id++;
   j->t_submit = s->Now();
   s->jobs[j->id] = j;
   s->pending.push_back(j);
   s->cv_work.notify_one();
   return j->id;
}

long long GmlxServeSubmit(size_t SrvIdx, int type, int n, unsigned long long seed, double param)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s || type < GmlxJobQuality || type > GmlxJobCG || n < 1 || n > 1000) return -1;
   return Enqueue(s, type, n, seed, param);
}

long long GmlxServeAddMesh(size_t SrvIdx, int NmbVer, int NmbTet, const float *crd, const int *ref, const int *tet)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s || NmbVer < 1 || NmbTet < 1 || !crd || !tet) return -1;
   for(long long i = 0; i < (long long)NmbTet * 4; i++)
      if(tet[i] < 0 || tet[i] >= NmbVer) return -1;
   auto m = std::make_shared<UserMesh>();
   m->nver = NmbVer;
   m->ntet = NmbTet;
   m->crd.assign(crd, crd + (size_t)NmbVer * 4);
   for(long long v = 0; v < NmbVer; v++) m->crd[v * 4 + 3] = 0.f;
   m->ref.assign((size_t)NmbVer, 0);
   if(ref) m->ref.assign(ref, ref + NmbVer);
   m->tet.assign(tet, tet + (size_t)NmbTet * 4);
   std::lock_guard<std::mutex> lk(s->mtx);
   const unsigned long long key = s->next_mesh++;
   m->seed = key;
   s->meshes[key] = m;
   return (long long)key;
}

int GmlxServeRemoveMesh(size_t SrvIdx, long long MeshKey)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s || MeshKey < 1) return 0;
   std::lock_guard<std::mutex> lk(s->mtx);
   const unsigned long long key = (unsigned long long)MeshKey;
   if(!s->meshes.count(key)) return 0;
   // refused while a job on the mesh is pending or running
   for(ServeJob *j : s->pending)
      if(j->n == 0 && j->seed == key) return 0;
   for(Resident *r : s->cache)
      if(r->n == 0 && r->seed == key && r->busy) return 0;
   for(auto it = s->cache.begin(); it != s->cache.end();)
   {
      Resident *r = *it;
      if(r->n == 0 && r->seed == key)
      {
         s->used -= r->bytes;
         Unload(r);
         delete r;
         it = s->cache.erase(it);
      }
      else ++it;
   }
   s->meshes.erase(key);
   return 1;
}

long long GmlxServeSubmitMesh(size_t SrvIdx, int type, long long MeshKey, double param)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s || type < GmlxJobQuality || type > GmlxJobCG || MeshKey < 1) return -1;
   {
      std::lock_guard<std::mutex> lk(s->mtx);
      if(!s->meshes.count((unsigned long long)MeshKey)) return -1;
   }
   return Enqueue(s, type, 0, (unsigned long long)MeshKey, param);
}

int GmlxServeWait(size_t SrvIdx, long long JobIdx, double *res, double *tim)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s) return -1;
   std::unique_lock<std::mutex> lk(s->mtx);
   auto it = s->jobs.find(JobIdx);
   if(it == s->jobs.end()) return -1;
   ServeJob *j = it->second;
   j->waiters++;
   s->nwaiters++;
   s->cv_done.wait(lk, [&]() { return j->done; });
   if(res) for(int i = 0; i < 4; i++) res[i] = j->res[i];
   if(tim)
   {
      tim[0] = j->t_submit;
      tim[1] = j->t_admit;
      tim[2] = j->t_done;
      tim[3] = j->loaded ? 1. : 0.;
   }
   const int rc = j->rc;
   // the first waiter to return retires the id; the last one frees the record
   if(!j->retired)
   {
      s->jobs.erase(JobIdx);
      j->retired = true;
   }
   s->nwaiters--;
   if(--j->waiters == 0) delete j;
   s->cv_done.notify_all();
   return rc;
}

int GmlxServeInfo(size_t SrvIdx, double *info)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s || !info) return 0;
   std::lock_guard<std::mutex> lk(s->mtx);
   const double v[8] = {(double)s->cache.size(), (double)s->used, (double)s->loads, (double)s->evictions,
                        (double)s->ndone, (double)s->max_running, (double)s->budget, (double)s->peak};
   for(int i = 0; i < 8; i++) info[i] = v[i];
   return 1;
}

void GmlxServeStop(size_t SrvIdx)
{
   Server *s = reinterpret_cast<Server *>(SrvIdx);
   if(!s) return;
   {
      std::lock_guard<std::mutex> lk(s->mtx);
      s->stop = true;
      // jobs not yet admitted end with -8; running ones finish
      for(ServeJob *j : s->pending)
      {
         j->t_admit = s->Now();
         Finish(s, j, -8);
      }
      s->pending.clear();
      s->cv_work.notify_all();
   }
   for(auto &t : s->workers) t.join();
   {
      // every job is done now: wake blocked waiters and let them leave
      std::unique_lock<std::mutex> lk(s->mtx);
      s->cv_done.notify_all();
      s->cv_done.wait(lk, [&]() { return s->nwaiters == 0; });
   }
   for(Resident *r : s->cache) { Unload(r); delete r; }
   for(auto &p : s->jobs) delete p.second;
   delete s;
}

} // extern "C"

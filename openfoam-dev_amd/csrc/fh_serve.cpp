// Request loop for many independent solves (fh_server_*, include/foamhip.h).
//
// Worker threads, one per solve slot, each take the oldest queued request
// once it is admissible, solve it with the ordinary fh_matrix_set / fh_solve
// path on a matrix that stays resident for its mesh, and queue the result.
// Every matrix owns its HIP stream, so solves of different slots overlap on
// the device; every reduction of a solve is per-solve and fixed-order, so a
// result does not depend on what ran beside it.
//
// Admission (in arrival order): a request runs when a slot is free and
// either an idle resident matrix of its mesh exists or the estimated
// footprint of a new one fits the memory budget -- after idle matrices of
// other meshes are released, least recently used first.  A request larger
// than the whole budget runs when nothing else is resident.
#include "fh_matrix.h"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace
{

struct Pending
{
    int64_t id = 0;
    fh_mesh* mesh = nullptr;
    std::vector<double> diag, upper, lower, source, psi;   // owned copies
    const double *pDiag = nullptr, *pUpper = nullptr, *pLower = nullptr, *pSource = nullptr;
    bool symmetric = true;
    fh_controls controls{};
    fh_controls precondControls{};     // owned copy of controls.preconditionerControls
    double submitTime = 0.0;
    int64_t footprint = 0;
};

struct Resident
{
    fh_matrix* matrix = nullptr;
    fh_mesh* mesh = nullptr;
    int64_t bytes = 0;
};

// device bytes of one resident matrix with its solver work space: FP64 and
// FP32 row values, face-ordered coefficients, packed FP32 tile
// coefficients, factors and about twenty work vectors
int64_t footprint_of(const fh_mesh* m)
{
    const int64_t N = m->nCells, F = m->nFaces, nnz = m->nnzPad;
    return 12*(N + nnz) + 16*F + 24*N + 40*N + 8*20*N + 4*12*N + (1 << 20);
}

} // namespace

struct fh_server
{
    fh_server_config cfg{};
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::unique_ptr<Pending>> queue;
    std::deque<fh_result> done;
    std::map<int64_t, std::vector<double>> solutions;
    std::list<Resident> idle;             // most recently used at the back
    int64_t residentBytes = 0;            // idle + in use
    int32_t inFlight = 0;
    int64_t pending = 0;                  // submitted, not yet returned by wait
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // under mu: pick up the head request if admissible; returns the resident
    // matrix to use (nullptr: a new one, its bytes already reserved)
    bool admit(std::unique_ptr<Pending>& job, Resident& res, std::vector<Resident>& release)
    {
        if (queue.empty() || inFlight >= cfg.maxConcurrent) return false;
        Pending* p = queue.front().get();
        for (auto it = idle.begin(); it != idle.end(); ++it)
        {
            if (it->mesh == p->mesh)
            {
                res = *it;
                idle.erase(it);
                job = std::move(queue.front());
                queue.pop_front();
                return true;
            }
        }
        // idle matrices (all of other meshes here) are released only when
        // the request is admitted with it: a refused request frees nothing,
        // so the bytes counted always match what is allocated
        const int64_t need = p->footprint;
        if (residentBytes + need > cfg.memoryBudget)
        {
            int64_t freeable = 0;
            for (const Resident& r : idle) freeable += r.bytes;
            if (residentBytes - freeable + need > cfg.memoryBudget && inFlight > 0)
                return false;              // wait for a running solve to end
            while (residentBytes + need > cfg.memoryBudget && !idle.empty())
            {
                release.push_back(idle.front());
                residentBytes -= idle.front().bytes;
                idle.pop_front();
            }
        }
        residentBytes += need;
        res = Resident{nullptr, p->mesh, need};
        job = std::move(queue.front());
        queue.pop_front();
        return true;
    }

    void worker()
    {
        for (;;)
        {
            std::unique_ptr<Pending> job;
            Resident res;
            std::vector<Resident> release;
            {
                std::unique_lock<std::mutex> lk(mu);
                cvWork.wait(lk, [&] { return stop || admit(job, res, release); });
                if (!job) return;
                inFlight++;
            }
            for (Resident& r : release) fh_matrix_destroy(r.matrix);
            (void)hipSetDevice(job->mesh->device);
            fh_result out;
            std::memset(&out, 0, sizeof(out));
            out.id = job->id;
            out.nCells = job->mesh->nCells;
            out.submitTime = job->submitTime;
            out.startTime = now();
            int st = FH_OK;
            if (!res.matrix) st = fh_matrix_create(job->mesh, &res.matrix);
            if (st == FH_OK)
                st = fh_matrix_set(res.matrix, job->pDiag, job->pUpper,
                                   job->symmetric ? nullptr : job->pLower);
            if (st == FH_OK)
                st = fh_solve(res.matrix, &job->controls, job->psi.data(), job->pSource,
                              &out.perf);
            out.endTime = now();
            out.status = st;
            if (st != FH_OK)
                std::snprintf(out.message, sizeof(out.message), "%s", fh_last_error());
            {
                std::lock_guard<std::mutex> lk(mu);
                inFlight--;
                if (res.matrix)
                    idle.push_back(res);
                else
                    residentBytes -= res.bytes;
                solutions[job->id] = std::move(job->psi);
                done.push_back(out);
            }
            cvDone.notify_all();
            cvWork.notify_all();
        }
    }
};

extern "C" {

void fh_server_config_default(fh_server_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->maxConcurrent = 8;
    c->memoryBudget = (int64_t)32 << 30;
}

int fh_server_create(const fh_server_config* config, fh_server** out)
{
    struct Ctx { const fh_server_config* c; fh_server** o; } ctx{config, out};
    return fh::guarded_call([](void* p) {
        Ctx* a = static_cast<Ctx*>(p);
        if (!a->o) fh::fail(FH_ERR_INVALID, "NULL output handle");
        *a->o = nullptr;
        fh_server_config cfg;
        fh_server_config_default(&cfg);
        if (a->c) cfg = *a->c;
        if (cfg.maxConcurrent < 1 || cfg.maxConcurrent > 64)
            fh::fail(FH_ERR_INVALID, "maxConcurrent must be in [1, 64] (got %d)",
                     cfg.maxConcurrent);
        if (cfg.memoryBudget <= 0) fh::fail(FH_ERR_INVALID, "memoryBudget must be positive");
        auto s = std::make_unique<fh_server>();
        s->cfg = cfg;
        fh_server* raw = s.get();
        for (int i = 0; i < cfg.maxConcurrent; i++)
            s->workers.emplace_back([raw] { raw->worker(); });
        *a->o = s.release();
    }, &ctx);
}

int fh_server_submit(fh_server* s, const fh_request* r)
{
    struct Ctx { fh_server* s; const fh_request* r; } ctx{s, r};
    return fh::guarded_call([](void* p) {
        Ctx* c = static_cast<Ctx*>(p);
        fh_server* s = c->s;
        const fh_request* r = c->r;
        if (!s || !r) fh::fail(FH_ERR_INVALID, "NULL argument");
        if (!r->mesh || !r->diag || !r->source) fh::fail(FH_ERR_INVALID, "request without mesh, diag or source");
        if (r->lower && !r->upper) fh::fail(FH_ERR_INVALID, "lower given without upper");
        const size_t N = r->mesh->nCells, F = r->mesh->nFaces;
        auto job = std::make_unique<Pending>();
        job->submitTime = s->now();
        job->id = r->id;
        job->mesh = r->mesh;
        job->symmetric = r->lower == nullptr;
        if (s->cfg.borrowInputs)
        {
            job->pDiag = r->diag;
            job->pUpper = r->upper;
            job->pLower = r->lower;
            job->pSource = r->source;
        }
        else
        {
            job->diag.assign(r->diag, r->diag + N);
            if (r->upper) job->upper.assign(r->upper, r->upper + F);
            if (r->lower) job->lower.assign(r->lower, r->lower + F);
            job->source.assign(r->source, r->source + N);
            job->pDiag = job->diag.data();
            job->pUpper = r->upper ? job->upper.data() : nullptr;
            job->pLower = r->lower ? job->lower.data() : nullptr;
            job->pSource = job->source.data();
        }
        if (r->psi0) job->psi.assign(r->psi0, r->psi0 + N);
        else job->psi.assign(N, 0.0);
        job->controls = r->controls;
        if (r->controls.preconditionerControls)
        {
            job->precondControls = *r->controls.preconditionerControls;
            job->precondControls.preconditionerControls = nullptr;
            job->controls.preconditionerControls = &job->precondControls;
        }
        job->footprint = footprint_of(r->mesh);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->solutions.count(r->id))
                fh::fail(FH_ERR_INVALID, "request id %lld already has an unfetched result",
                         (long long)r->id);
            s->queue.push_back(std::move(job));
            s->pending++;
        }
        s->cvWork.notify_all();
    }, &ctx);
}

int fh_server_wait(fh_server* s, double timeoutSeconds, fh_result* result, int32_t* got)
{
    struct Ctx { fh_server* s; double t; fh_result* r; int32_t* g; } ctx{s, timeoutSeconds, result, got};
    return fh::guarded_call([](void* p) {
        Ctx* c = static_cast<Ctx*>(p);
        if (!c->s || !c->r || !c->g) fh::fail(FH_ERR_INVALID, "NULL argument");
        *c->g = 0;
        std::unique_lock<std::mutex> lk(c->s->mu);
        const auto deadline = std::chrono::steady_clock::now()
                            + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                  std::chrono::duration<double>(c->t > 0 ? c->t : 0.0));
        c->s->cvDone.wait_until(lk, deadline, [&] { return !c->s->done.empty(); });
        if (c->s->done.empty()) return;
        *c->r = c->s->done.front();
        c->s->done.pop_front();
        c->s->pending--;
        *c->g = 1;
    }, &ctx);
}

int fh_server_fetch(fh_server* s, int64_t id, double* psi)
{
    struct Ctx { fh_server* s; int64_t id; double* psi; } ctx{s, id, psi};
    return fh::guarded_call([](void* p) {
        Ctx* c = static_cast<Ctx*>(p);
        if (!c->s || !c->psi) fh::fail(FH_ERR_INVALID, "NULL argument");
        std::lock_guard<std::mutex> lk(c->s->mu);
        auto it = c->s->solutions.find(c->id);
        if (it == c->s->solutions.end())
            fh::fail(FH_ERR_INVALID, "no completed result for request %lld", (long long)c->id);
        std::memcpy(c->psi, it->second.data(), it->second.size()*sizeof(double));
        c->s->solutions.erase(it);
    }, &ctx);
}

int fh_server_now(fh_server* s, double* seconds)
{
    if (!s || !seconds) return FH_ERR_INVALID;
    *seconds = s->now();
    return FH_OK;
}

int fh_server_pending(fh_server* s, int64_t* count)
{
    if (!s || !count) return FH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *count = s->pending;
    return FH_OK;
}

int fh_server_destroy(fh_server* s)
{
    if (!s) return FH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cvWork.notify_all();
    for (auto& t : s->workers) t.join();
    for (Resident& r : s->idle) fh_matrix_destroy(r.matrix);
    delete s;
    return FH_OK;
}

} // extern "C"

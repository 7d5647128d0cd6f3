// Request loop over device-resident engines: admission by HBM budget, factor
// reuse keyed by (mesh, frequency, mode), overlapping requests on one GPU.
// Host-side C++ only; the GPU work is the engines' (one HIP stream each).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/peec_serve.h"
#include "common.h"

namespace {

using Clock = std::chrono::steady_clock;

struct MeshEntry {
    const Mesh *mesh = nullptr;
    const DualMesh *dual = nullptr;
    size_t ne = 0, nv = 0;
    std::vector<double> L, P;  // static matrices, for engines after the first
    size_t bytes = 0;          // estimated working set of one engine
};

struct EngineEntry {
    int id = 0;
    int mesh = -1;
    PeecEngine *eng = nullptr;
    size_t bytes = 0;
    bool busy = false;
    bool ready = false;  // created (a reserved entry is being created)
    bool has_factor = false;
    double f = 0.0;
    int mode = -1;
    long long last_use = 0;
};

struct Job {
    long long ticket = 0;
    int mesh = 0;
    double f = 0.0;
    int mode = 0;
    std::vector<int> ports;
    std::vector<Complex> z;
    PeecRequestResult result{};
    Clock::time_point t_submit, t_start;
    bool done = false;
};

// Device bytes of one engine of a mesh: static Lp/P (FP64) and the
// per-point working set.
size_t engine_bytes(size_t ne, size_t nv, bool retarded) {
    return 8 * (ne * ne + nv * nv) + peec::engine_point_bytes(ne, nv, retarded);
}

// Further device bytes of each batch member beyond the first: its FP32
// factor, retarded Lp(w)/P(w), scales and operand packs.
size_t member_bytes(size_t ne, size_t nv, bool retarded) {
    const size_t n = ne + nv;
    return 8 * n * n + (retarded ? 16 * (ne * ne + nv * nv) : 0) + 64 * n;
}

}  // namespace

struct PeecServer {
    PeecServerOptions opt{};
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::unique_ptr<MeshEntry>> meshes;
    std::vector<std::unique_ptr<EngineEntry>> engines;
    std::deque<Job *> queue;
    std::unordered_map<long long, std::unique_ptr<Job>> jobs;
    std::vector<std::thread> workers;
    long long next_ticket = 1, clock = 0;
    int next_engine = 0;
    bool stop = false;

    void worker();
    Job *pick(EngineEntry *&e, std::vector<PeecEngine *> &evicted);
    void gather(Job *j, EngineEntry *e, std::vector<Job *> &batch);
};

// Under the lock: the mixed requests of j's mesh waiting in the queue join
// j's batch on engine e (up to max_batch in all), each as long as the
// engine's grown working set fits under the budget. A request that finds
// its factor held by e runs alone.
void PeecServer::gather(Job *j, EngineEntry *e, std::vector<Job *> &batch) {
    batch.assign(1, j);
    if (j->mode != PEEC_SOLVE_MIXED || opt.max_batch <= 1) return;
    if (e->has_factor && e->f == j->f && e->mode == j->mode) return;
    const MeshEntry &me = *meshes[j->mesh];
    const size_t extra = member_bytes(me.ne, me.nv, opt.engine.retarded != 0);
    for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < opt.max_batch;) {
        Job *o = *it;
        if (o->mesh != j->mesh || o->mode != PEEC_SOLVE_MIXED) {
            ++it;
            continue;
        }
        const size_t need = me.bytes + batch.size() * extra;
        if (need > e->bytes) {
            if (used - e->bytes + need > budget) break;
            used += need - e->bytes;
            e->bytes = need;
        }
        batch.push_back(o);
        it = queue.erase(it);
    }
}

// Under the lock: the first queued request that can run now, and its engine
// (an idle one holding the request's factor, an idle one of its mesh, or a
// new reserved entry that fits the budget once idle engines of other meshes
// are evicted, least recently used first). Requests that cannot run yet are
// passed over.
Job *PeecServer::pick(EngineEntry *&e, std::vector<PeecEngine *> &evicted) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
        Job *j = *it;
        EngineEntry *same = nullptr, *cached = nullptr;
        for (auto &en : engines) {
            if (en->busy || !en->ready || en->mesh != j->mesh) continue;
            if (en->has_factor && en->f == j->f && en->mode == j->mode) cached = en.get();
            if (!same || en->last_use < same->last_use) same = en.get();
        }
        EngineEntry *take = cached ? cached : same;
        if (!take) {
            const size_t need = meshes[j->mesh]->bytes;
            if (used + need > budget) {
                // evict idle engines of other meshes, least recently used first
                std::vector<EngineEntry *> idle;
                for (auto &en : engines)
                    if (!en->busy && en->ready && en->mesh != j->mesh) idle.push_back(en.get());
                std::sort(idle.begin(), idle.end(),
                          [](const EngineEntry *a, const EngineEntry *b) { return a->last_use < b->last_use; });
                size_t freeable = 0;
                for (EngineEntry *x : idle) freeable += x->bytes;
                if (used - freeable + need > budget) continue;  // cannot fit yet
                for (EngineEntry *x : idle) {
                    if (used + need <= budget) break;
                    used -= x->bytes;
                    evicted.push_back(x->eng);
                    engines.erase(std::find_if(engines.begin(), engines.end(),
                                               [x](const std::unique_ptr<EngineEntry> &p) { return p.get() == x; }));
                }
            }
            auto ne = std::make_unique<EngineEntry>();
            ne->id = next_engine++;
            ne->mesh = j->mesh;
            ne->bytes = need;
            used += need;
            take = ne.get();
            engines.push_back(std::move(ne));
        }
        take->busy = true;
        take->last_use = ++clock;
        queue.erase(it);
        e = take;
        return j;
    }
    return nullptr;
}

void PeecServer::worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        EngineEntry *e = nullptr;
        std::vector<PeecEngine *> evicted;
        Job *j = nullptr;
        cv_work.wait(lk, [&] {
            if (stop && queue.empty()) return true;
            j = pick(e, evicted);
            return j != nullptr;
        });
        if (!j) return;
        std::vector<Job *> batch;
        gather(j, e, batch);
        lk.unlock();
        for (PeecEngine *x : evicted) peec_engine_destroy(x);
        const Clock::time_point t_start = Clock::now();
        for (Job *b : batch) b->t_start = t_start;
        int status = 0;
        if (!e->ready) {
            const MeshEntry &m = *meshes[j->mesh];
            PeecEngine *eng = nullptr;
            status = peec_engine_create_with_matrices(m.mesh, m.dual, &opt.engine, m.L.data(), m.P.data(), &eng);
            e->eng = eng;
        }
        const size_t nb = batch.size();
        std::vector<PeecPointReport> rep(nb);
        std::vector<int> st(nb, status);
        if (status == 0 && nb == 1) {
            st[0] = peec_engine_solve(e->eng, j->f, j->mode, j->ports.data(), j->ports.size() / 2, j->z.data(),
                                      nullptr, nullptr, &rep[0]);
        } else if (status == 0) {
            std::vector<double> f(nb);
            std::vector<const int *> ports(nb);
            std::vector<size_t> nports(nb);
            std::vector<Complex *> z(nb);
            for (size_t i = 0; i < nb; ++i) {
                f[i] = batch[i]->f;
                ports[i] = batch[i]->ports.data();
                nports[i] = batch[i]->ports.size() / 2;
                z[i] = batch[i]->z.data();
            }
            peec_engine_solve_batch(e->eng, nb, f.data(), ports.data(), nports.data(), z.data(), rep.data(),
                                    st.data());
        }
        const Clock::time_point t_end = Clock::now();
        lk.lock();
        int engine_id = 0;
        if (!e->ready && !e->eng) {
            // creation failed: release the reservation
            used -= e->bytes;
            engines.erase(std::find_if(engines.begin(), engines.end(),
                                       [e](const std::unique_ptr<EngineEntry> &p) { return p.get() == e; }));
        } else {
            e->ready = true;
            e->busy = false;
            // a batch leaves no single point's factor behind
            e->has_factor = nb == 1 && st[0] == 0 && !rep[0].used_fp64_fallback;
            e->f = j->f;
            e->mode = j->mode;
            engine_id = e->id;
        }
        for (size_t i = 0; i < nb; ++i) {
            Job *b = batch[i];
            b->result.engine = engine_id;
            b->result.status = st[i];
            b->result.report = rep[i];
            b->result.queue_s = std::chrono::duration<double>(b->t_start - b->t_submit).count();
            b->result.latency_s = std::chrono::duration<double>(t_end - b->t_submit).count();
            b->done = true;
        }
        cv_done.notify_all();
        cv_work.notify_all();  // an engine is free again
    }
}

extern "C" {

void peec_server_options_default(PeecServerOptions *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    peec_engine_options_default(&o->engine);
    o->slots = 4;
    o->hbm_budget_bytes = 0;
    o->max_batch = 8;
}

int peec_server_create(const PeecServerOptions *o, PeecServer **out) {
    if (!o || !out || o->slots < 1 || o->slots > 64 || o->max_batch < 1 || o->max_batch > 64) return -1;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        PEEC_ERR("no usable GPU for the request server.");
        return -1;
    }
    auto sv = std::make_unique<PeecServer>();
    sv->opt = *o;
    // the admission estimate has no room for stored moments of the smooth
    // remainder: engines assemble every point by quadrature
    sv->opt.engine.smooth_moments = 0;
    sv->opt.engine.shifted_moments = 0;
    sv->budget = o->hbm_budget_bytes ? o->hbm_budget_bytes : free_b / 10 * 9;
    for (int i = 0; i < o->slots; ++i) sv->workers.emplace_back([s = sv.get()] { s->worker(); });
    *out = sv.release();
    return 0;
}

int peec_server_add_mesh(PeecServer *sv, const Mesh *mesh, const DualMesh *dual, int *mesh_id) {
    if (!sv || !mesh || !dual || !mesh_id) return -1;
    auto m = std::make_unique<MeshEntry>();
    m->mesh = mesh;
    m->dual = dual;
    m->ne = mesh->n_edges;
    m->nv = mesh->n_nodes;
    m->bytes = engine_bytes(m->ne, m->nv, sv->opt.engine.retarded != 0);
    std::vector<PeecEngine *> evicted;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (m->bytes > sv->budget) {
            PEEC_ERR("one engine of this mesh needs %zu bytes, more than the budget of %zu.", m->bytes, sv->budget);
            return -1;
        }
        // room for the mesh's first engine: idle engines go, least recently
        // used first (busy ones stay; the first engine is created regardless)
        while (sv->used + m->bytes > sv->budget) {
            EngineEntry *lru = nullptr;
            for (auto &en : sv->engines)
                if (!en->busy && en->ready && (!lru || en->last_use < lru->last_use)) lru = en.get();
            if (!lru) break;
            sv->used -= lru->bytes;
            evicted.push_back(lru->eng);
            sv->engines.erase(std::find_if(sv->engines.begin(), sv->engines.end(),
                                           [lru](const std::unique_ptr<EngineEntry> &p) { return p.get() == lru; }));
        }
    }
    for (PeecEngine *x : evicted) peec_engine_destroy(x);
    PeecEngine *eng = nullptr;
    if (peec_engine_create(mesh, dual, &sv->opt.engine, &eng) != 0) return -1;
    m->L.resize(m->ne * m->ne);
    m->P.resize(m->nv * m->nv);
    if (peec_engine_static_matrices(eng, m->L.data(), m->P.data()) != 0) {
        peec_engine_destroy(eng);
        return -1;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    const int id = (int)sv->meshes.size();
    auto e = std::make_unique<EngineEntry>();
    e->id = sv->next_engine++;
    e->mesh = id;
    e->eng = eng;
    e->bytes = m->bytes;
    e->ready = true;
    e->last_use = ++sv->clock;
    sv->used += m->bytes;
    sv->engines.push_back(std::move(e));
    sv->meshes.push_back(std::move(m));
    *mesh_id = id;
    sv->cv_work.notify_all();
    return 0;
}

int peec_server_submit(PeecServer *sv, int mesh_id, double frequency, int mode, const int *ports, size_t n_ports,
                       long long *ticket) {
    if (!sv || !ports || n_ports == 0 || !ticket || !(frequency > 0.0)) return -1;
    if (mode != PEEC_SOLVE_FP64 && mode != PEEC_SOLVE_MIXED) return -1;
    auto j = std::make_unique<Job>();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (mesh_id < 0 || mesh_id >= (int)sv->meshes.size()) return -1;
        const size_t nv = sv->meshes[mesh_id]->nv;
        for (size_t k = 0; k < 2 * n_ports; ++k)
            if (ports[k] < 0 || (size_t)ports[k] >= nv) return -1;
        j->ticket = sv->next_ticket++;
        j->mesh = mesh_id;
        j->f = frequency;
        j->mode = mode;
        j->ports.assign(ports, ports + 2 * n_ports);
        j->z.assign(n_ports, Complex{0.0, 0.0});
        j->t_submit = Clock::now();
        *ticket = j->ticket;
        sv->queue.push_back(j.get());
        sv->jobs.emplace(j->ticket, std::move(j));
    }
    sv->cv_work.notify_one();
    return 0;
}

int peec_server_wait(PeecServer *sv, long long ticket, Complex *z_in, PeecRequestResult *result) {
    if (!sv) return -1;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) return -1;
    Job *j = it->second.get();
    sv->cv_done.wait(lk, [j] { return j->done; });
    if (z_in) std::memcpy(z_in, j->z.data(), sizeof(Complex) * j->z.size());
    if (result) *result = j->result;
    const int status = j->result.status;
    sv->jobs.erase(it);
    return status;
}

size_t peec_server_engine_bytes(size_t n_edges, size_t n_nodes, int retarded) {
    return engine_bytes(n_edges, n_nodes, retarded != 0);
}

size_t peec_server_memory(const PeecServer *sv, int *n_engines) {
    if (!sv) return 0;
    auto *s = const_cast<PeecServer *>(sv);
    std::lock_guard<std::mutex> lk(s->mu);
    if (n_engines) *n_engines = (int)s->engines.size();
    return s->used;
}

void peec_server_destroy(PeecServer *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (std::thread &t : sv->workers) t.join();
    for (auto &e : sv->engines)
        if (e->eng) peec_engine_destroy(e->eng);
    delete sv;
}

}  // extern "C"

// lq_server.cpp -- continuous batching of independent runs on one engine.
//
// The reference runs one simulation per process (src/loquat.cu:216-475);
// this layer serves a stream of them.  Requests queue in arrival order; at
// every tick the head of the queue is admitted while its device footprint
// (lq_engine_bytes_per_run) fits the HBM budget and the engine's particle,
// cell and run capacities, then every resident run advances steps_per_tick
// steps in the same launches, and finished or failed runs are evicted
// (optionally downloading their state).  Admission is strict FIFO: a large
// request at the head is never overtaken, so no request starves.
//
// Built only on the engine's C ABI.  With fake_backend = 1 no GPU is touched:
// runs advance steps_per_tick steps per tick on the host, and a run whose
// input holds a non-finite value fails at its first tick -- the same
// contract the device path keeps through lq_engine_poll.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/loquat_amd.h"

namespace lq {
void set_error(const std::string& msg);
lq_status check_run_input(const lq_run_input* in, lq_params* par_out);
}

namespace {

struct Request {
    int32_t id = -1;
    lq_params par{};
    lq_vw_params vw{};
    std::vector<float> pospres, velrhop, idep, str, stn;
    std::vector<double> posd;
    int64_t nsteps = 0;
    int64_t bytes = 0;
    int32_t ncell = 0;
    int32_t state = LQ_REQ_QUEUED;
    std::string why;         // reason of a rejection
    int32_t run = -1;        // engine run id while running
    int64_t steps_done = 0;
    bool bad_input = false;  // fake backend: non-finite input
    bool has_result = false;
    std::vector<float> r_pospres, r_velrhop, r_idep, r_str, r_stn;
    std::vector<double> r_posd;
};

}  // namespace

struct lq_server {
    lq_server_config cfg{};
    lq_engine* eng = nullptr;
    std::map<int32_t, Request> reqs;
    std::deque<int32_t> queue;
    std::vector<int32_t> running;
    int64_t bytes_in_use = 0, np_in_use = 0, ncell_in_use = 0, npm_in_use = 0;
    int32_t next_id = 0;
};

static bool finite_all(const float* a, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

extern "C" lq_status lq_server_create(const lq_server_config* cfg, lq_server** out) {
    if (!cfg || !out || cfg->np_cap <= 0 || cfg->ncell_cap <= 0 || cfg->max_runs <= 0 || cfg->hbm_budget <= 0 ||
        cfg->steps_per_tick <= 0) {
        lq::set_error("lq_server_create: invalid configuration");
        return LQ_ERR_INVALID_ARG;
    }
    lq_server* s = new lq_server();
    s->cfg = *cfg;
    if (!cfg->fake_backend) {
        lq_status st = lq_engine_create(cfg->np_cap, cfg->ncell_cap, cfg->max_runs, &s->eng);
        if (st != LQ_OK) {
            delete s;
            return st;
        }
    }
    *out = s;
    return LQ_OK;
}

extern "C" lq_status lq_server_destroy(lq_server* s) {
    if (!s) return LQ_OK;
    if (s->eng) lq_engine_destroy(s->eng);
    delete s;
    return LQ_OK;
}

extern "C" lq_status lq_server_submit(lq_server* s, const lq_run_input* in, int32_t* request_id) {
    if (!s || !in || !request_id || !in->host.pospres || !in->host.velrhop || !in->host.idep || !in->host.str ||
        !in->host.stn) {
        lq::set_error("lq_server_submit: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    lq_params par = in->par;
    if (lq_params_derive(&par) != LQ_OK || par.np < 0 || par.np != par.npb + par.npm) {
        lq::set_error("lq_server_submit: np != npb + npm or parameters that cannot be derived");
        return LQ_ERR_INVALID_ARG;
    }
    // the checks the engine makes when the run is admitted, made now: a
    // request that would fail them is rejected here instead of blocking the
    // head of the queue
    lq_params checked;
    const lq_status valid = lq::check_run_input(in, &checked);
    Request r;
    r.par = in->par;
    r.vw = in->vw;
    const size_t n = (size_t)par.np;
    r.pospres.assign(in->host.pospres, in->host.pospres + 4 * n);
    r.velrhop.assign(in->host.velrhop, in->host.velrhop + 4 * n);
    r.idep.assign(in->host.idep, in->host.idep + 4 * n);
    r.str.assign(in->host.str, in->host.str + 6 * n);
    r.stn.assign(in->host.stn, in->host.stn + 6 * n);
    if (in->posd) r.posd.assign(in->posd, in->posd + 3 * n);
    r.ncell = par.ncell;
    r.bytes = lq_engine_bytes_per_run(par.np, par.ncell);
    if (in->nsteps > 0) {
        r.nsteps = in->nsteps;
    } else {
        double t = (double)par.time;
        int64_t k = 0;
        while (t < (double)par.time_max) {
            t += (double)par.dt;
            ++k;
        }
        r.nsteps = k;
    }
    r.bad_input = !(finite_all(r.pospres.data(), r.pospres.size()) && finite_all(r.velrhop.data(), r.velrhop.size()) &&
                    finite_all(r.str.data(), r.str.size()));
    r.id = s->next_id++;
    const lq_server_config& c = s->cfg;
    if (valid != LQ_OK) {
        r.state = LQ_REQ_REJECTED;
        r.why = lq_last_error();
    } else if (r.bytes > c.hbm_budget || par.np > c.np_cap || par.ncell > c.ncell_cap) {
        r.state = LQ_REQ_REJECTED;
        r.why = "larger than the HBM budget or the engine capacities";
    } else {
        r.state = LQ_REQ_QUEUED;
        s->queue.push_back(r.id);
    }
    if (r.state == LQ_REQ_REJECTED) {
        std::vector<float>().swap(r.pospres);
        std::vector<float>().swap(r.velrhop);
        std::vector<float>().swap(r.idep);
        std::vector<float>().swap(r.str);
        std::vector<float>().swap(r.stn);
        std::vector<double>().swap(r.posd);
    }
    *request_id = r.id;
    s->reqs[r.id] = std::move(r);
    return LQ_OK;
}

static lq_status admit(lq_server* s, Request& r) {
    if (s->eng) {
        lq_run_input in{};
        in.par = r.par;
        in.vw = r.vw;
        in.host.pospres = r.pospres.data();
        in.host.velrhop = r.velrhop.data();
        in.host.idep = r.idep.data();
        in.host.str = r.str.data();
        in.host.stn = r.stn.data();
        in.posd = r.posd.empty() ? nullptr : r.posd.data();
        in.nsteps = r.nsteps;
        lq_status st = lq_engine_add_run(s->eng, &in, &r.run);
        if (st != LQ_OK) return st;
    }
    r.state = LQ_REQ_RUNNING;
    s->running.push_back(r.id);
    s->bytes_in_use += r.bytes;
    s->np_in_use += r.par.np;
    s->npm_in_use += r.par.npm;
    s->ncell_in_use += r.ncell;
    return LQ_OK;
}

static lq_status evict(lq_server* s, Request& r, int32_t final_state) {
    if (s->eng) {
        if (final_state == LQ_REQ_DONE && s->cfg.keep_results) {
            const size_t n = (size_t)r.par.np;
            r.r_pospres.resize(4 * n);
            r.r_velrhop.resize(4 * n);
            r.r_idep.resize(4 * n);
            r.r_str.resize(6 * n);
            r.r_stn.resize(6 * n);
            r.r_posd.resize(3 * n);
            lq_host_particles out = {r.r_pospres.data(), r.r_velrhop.data(), r.r_idep.data(), r.r_str.data(),
                                     r.r_stn.data()};
            lq_status st = lq_engine_download(s->eng, r.run, &out, r.r_posd.data());
            if (st != LQ_OK) return st;
            r.has_result = true;
        }
        lq_status st = lq_engine_remove_run(s->eng, r.run);
        if (st != LQ_OK) return st;
    }
    r.state = final_state;
    r.run = -1;
    s->bytes_in_use -= r.bytes;
    s->np_in_use -= r.par.np;
    s->npm_in_use -= r.par.npm;
    s->ncell_in_use -= r.ncell;
    // input copies are no longer needed
    std::vector<float>().swap(r.pospres);
    std::vector<float>().swap(r.velrhop);
    std::vector<float>().swap(r.idep);
    std::vector<float>().swap(r.str);
    std::vector<float>().swap(r.stn);
    std::vector<double>().swap(r.posd);
    return LQ_OK;
}

extern "C" lq_status lq_server_tick(lq_server* s, int32_t* finished_ids, int32_t cap, int32_t* nfinished) {
    if (!s || (cap > 0 && !finished_ids) || cap < 0) {
        lq::set_error("lq_server_tick: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    const lq_server_config& c = s->cfg;
    // 1. admission, strict FIFO.  A request the engine refuses (submission
    // already ran the engine's checks, so this is rare: a device error) is
    // rejected and dropped from the queue, so it cannot block the requests
    // behind it; a device error is still returned once it is dropped.
    int32_t nout = 0;
    lq_status admit_err = LQ_OK;
    while (!s->queue.empty()) {
        Request& r = s->reqs[s->queue.front()];
        const bool fits = s->bytes_in_use + r.bytes <= c.hbm_budget && s->np_in_use + r.par.np <= c.np_cap &&
                          s->ncell_in_use + r.ncell <= c.ncell_cap && (int32_t)s->running.size() < c.max_runs;
        if (!fits) break;
        lq_status st = admit(s, r);
        s->queue.pop_front();
        if (st != LQ_OK) {
            r.state = LQ_REQ_REJECTED;
            r.why = lq_last_error();
            if (nout < cap) finished_ids[nout] = r.id;
            ++nout;
            if (st == LQ_ERR_HIP || st == LQ_ERR_OUT_OF_MEMORY) {
                admit_err = st;
                break;
            }
        }
    }
    if (admit_err != LQ_OK) {
        if (nfinished) *nfinished = nout;
        return admit_err;
    }
    // 2. one chunk of steps for every resident run
    if (!s->running.empty()) {
        if (s->eng) {
            lq_status st = lq_engine_step(s->eng, c.steps_per_tick);
            if (st != LQ_OK) return st;
            int32_t nf = 0;
            st = lq_engine_poll(s->eng, &nf);
            if (st != LQ_OK) return st;
        }
    }
    // 3. eviction of finished / failed runs
    std::vector<int32_t> still;
    for (int32_t id : s->running) {
        Request& r = s->reqs[id];
        int32_t state = LQ_REQ_RUNNING;
        if (s->eng) {
            int64_t steps = 0;
            double t = 0;
            int32_t fin = 0, np = 0, failed = 0;
            lq_status st = lq_engine_run_info(s->eng, r.run, &steps, &t, &fin, &np);
            if (st == LQ_OK) st = lq_engine_run_failed(s->eng, r.run, &failed);
            if (st != LQ_OK) return st;
            r.steps_done = steps;
            if (failed)
                state = LQ_REQ_FAILED;
            else if (fin)
                state = LQ_REQ_DONE;
        } else {
            if (r.bad_input) {
                state = LQ_REQ_FAILED;
            } else {
                r.steps_done = std::min<int64_t>(r.nsteps, r.steps_done + c.steps_per_tick);
                if (r.steps_done >= r.nsteps) state = LQ_REQ_DONE;
            }
        }
        if (state == LQ_REQ_RUNNING) {
            still.push_back(id);
            continue;
        }
        lq_status st = evict(s, r, state);
        if (st != LQ_OK) return st;
        if (nout < cap) finished_ids[nout] = id;
        ++nout;
    }
    s->running.swap(still);
    if (nfinished) *nfinished = nout;
    return LQ_OK;
}

extern "C" lq_status lq_server_request_info(const lq_server* s, int32_t id, int32_t* state, int64_t* steps_done,
                                            int64_t* nsteps, int64_t* bytes) {
    if (!s) {
        lq::set_error("null server");
        return LQ_ERR_INVALID_ARG;
    }
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
        lq::set_error("unknown request id");
        return LQ_ERR_INVALID_ARG;
    }
    const Request& r = it->second;
    if (state) *state = r.state;
    if (steps_done) *steps_done = r.steps_done;
    if (nsteps) *nsteps = r.nsteps;
    if (bytes) *bytes = r.bytes;
    return LQ_OK;
}

extern "C" lq_status lq_server_result(lq_server* s, int32_t id, lq_host_particles* out, double* posd) {
    if (!s || !out) {
        lq::set_error("lq_server_result: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second.has_result) {
        lq::set_error("no result for this request (not done, failed, or keep_results = 0)");
        return LQ_ERR_STATE;
    }
    Request& r = it->second;
    const size_t n = (size_t)r.par.np;
    if (out->pospres) std::memcpy(out->pospres, r.r_pospres.data(), 16 * n);
    if (out->velrhop) std::memcpy(out->velrhop, r.r_velrhop.data(), 16 * n);
    if (out->idep) std::memcpy(out->idep, r.r_idep.data(), 16 * n);
    if (out->str) std::memcpy(out->str, r.r_str.data(), 24 * n);
    if (out->stn) std::memcpy(out->stn, r.r_stn.data(), 24 * n);
    if (posd) std::memcpy(posd, r.r_posd.data(), 24 * n);
    return LQ_OK;
}

extern "C" lq_status lq_server_stats(const lq_server* s, int32_t* queued, int32_t* running, int64_t* bytes_in_use,
                                     int64_t* npm_running) {
    if (!s) {
        lq::set_error("null server");
        return LQ_ERR_INVALID_ARG;
    }
    if (queued) *queued = (int32_t)s->queue.size();
    if (running) *running = (int32_t)s->running.size();
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
    if (npm_running) *npm_running = s->npm_in_use;
    return LQ_OK;
}

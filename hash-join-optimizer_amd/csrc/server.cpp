// server.cpp - queries over resident tables served by a pool of contexts on one GPU.
//
// No counterpart in the reference (its execute() runs one query at a time). Each slot is a
// context (its own stream, HBM arena and pinned staging) driven by its own host thread; queries
// wait in one FIFO queue and the next one is admitted as soon as a slot is free and the memory it
// reserves (an estimate of its staging, build tables and output) fits what the running queries
// leave of the server's budget, so queries that arrive while others run overlap on the device. A
// query whose reservation exceeds a slot's arena or the whole budget is refused with HJ_E_BUSY.
// The slots share one directory cache: a dimension's table, built once for a (table, key,
// filters) triple, is probed from HBM by every later query that uses it.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/hjoin.h"
#include "hj_internal.hpp"

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Job {
    uint64_t ticket;
    uint64_t reserve = 0;
    const hj_plan* plan;
    std::vector<const hj_dtable*> tables;
    uint64_t submit_ns = 0, start_ns = 0, done_ns = 0;
    int rc = 0;
    std::string error;
    hj_result result{};
    bool done = false;
};

}  // namespace

struct hj_server {
    std::vector<hj_context*> ctx;
    std::vector<std::thread> workers;
    uint64_t arena_bytes = 0;
    uint64_t budget = 0;    // bytes the running queries may reserve together
    uint64_t reserved = 0;  // what they reserve now
    std::shared_ptr<hj::DirCache> dcache;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::map<uint64_t, Job*> jobs;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::string last_error;
};

namespace {

// bytes a query reserves in a slot's arena: per scan, its rows' staged tuples, a build table
// (4 x the rows at 8 B per key slot) and row ids; per output column, the source rows' values
uint64_t reserve_bytes(const hj_plan* p, const hj_dtable* const* t) {
    uint64_t scans = 0, widest = 0;
    for (uint64_t i = 0; i < p->num_nodes; ++i) {
        const hj_node& nd = p->nodes[i];
        if (nd.kind != HJ_NODE_SCAN) continue;
        const uint64_t rows = hj_table_num_rows(t[nd.base_table_id]);
        scans += rows * (8 + 32 + 8);
        widest = rows > widest ? rows : widest;
    }
    const hj_node& root = p->nodes[p->root];
    return scans + widest * 16 * (root.num_output_attrs + 1);
}

void worker(hj_server* s, hj_context* ctx) {
    for (;;) {
        Job* j = nullptr;
        {
            // FIFO: the queue's first query runs once its reservation fits what the running
            // queries leave of the budget
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [s] {
                return (s->stop && s->queue.empty()) ||
                       (!s->queue.empty() && s->reserved + s->queue.front()->reserve <= s->budget);
            });
            if (s->queue.empty()) return;
            j = s->queue.front();
            s->queue.pop_front();
            s->reserved += j->reserve;
        }
        j->start_ns = now_ns();
        hj_dresult* r = nullptr;
        int rc = hj_execute_device(ctx, j->plan, j->tables.data(), &r);
        if (rc == HJ_OK) {
            rc = hj_dresult_to_host(ctx, r, &j->result);
            hj_dresult_free(ctx, r);
        }
        if (rc != HJ_OK) j->error = hj_last_error(ctx);
        j->rc = rc;
        j->done_ns = now_ns();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->done = true;
            s->reserved -= j->reserve;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // the memory it reserved may admit the next query
    }
}

}  // namespace

extern "C" {

int hj_server_create(uint32_t slots, uint64_t arena_bytes_per_slot, hj_server** out) {
    if (!out || !slots || slots > 64) return HJ_E_INVALID;
    *out = nullptr;
    auto* s = new hj_server();
    s->arena_bytes = arena_bytes_per_slot ? arena_bytes_per_slot : (4ull << 30);
    s->budget = s->arena_bytes * slots;
    s->dcache = std::make_shared<hj::DirCache>();
    s->dcache->budget = 1ull << 30;
    for (uint32_t i = 0; i < slots; ++i) {
        hj_context* c = nullptr;
        const int rc = hj_context_create(s->arena_bytes, &c);
        if (rc) {
            for (auto* x : s->ctx) hj_context_destroy(x);
            delete s;
            return rc;
        }
        c->dcache = s->dcache;
        s->ctx.push_back(c);
    }
    for (auto* c : s->ctx) s->workers.emplace_back(worker, s, c);
    *out = s;
    return HJ_OK;
}

int hj_server_destroy(hj_server* s) {
    if (!s) return HJ_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& w : s->workers) w.join();
    for (auto& kv : s->jobs) {
        hj_result_free(&kv.second->result);
        delete kv.second;
    }
    for (auto* c : s->ctx) hj_context_destroy(c);
    delete s;
    return HJ_OK;
}

int hj_server_submit(hj_server* s, const hj_plan* plan, const hj_dtable* const* tables,
                     uint64_t* ticket) {
    if (!s || !plan || !ticket || (!tables && plan->num_inputs) || !plan->nodes ||
        plan->root >= plan->num_nodes)
        return HJ_E_INVALID;
    for (uint64_t i = 0; i < plan->num_nodes; ++i)
        if (plan->nodes[i].kind == HJ_NODE_SCAN && plan->nodes[i].base_table_id >= plan->num_inputs)
            return HJ_E_INVALID;
    const uint64_t reserve = reserve_bytes(plan, tables);
    if (reserve > s->arena_bytes) return HJ_E_BUSY;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (reserve > s->budget) return HJ_E_BUSY;  // could never be admitted
    }
    auto* j = new Job();
    j->reserve = reserve;
    j->plan = plan;
    j->tables.assign(tables, tables + plan->num_inputs);
    j->submit_ns = now_ns();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->ticket = s->next_ticket++;
        s->jobs[j->ticket] = j;
        s->queue.push_back(j);
    }
    s->cv_work.notify_one();
    *ticket = j->ticket;
    return HJ_OK;
}

int hj_server_wait(hj_server* s, uint64_t ticket, hj_result* result, uint64_t* queued_ns,
                   uint64_t* run_ns) {
    if (!s || !result) return HJ_E_INVALID;
    std::memset(result, 0, sizeof *result);
    Job* j = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end()) return HJ_E_INVALID;  // unknown, or claimed by another waiter
        j = it->second;
        s->jobs.erase(it);  // claimed: this waiter alone owns the job from here on
        s->cv_done.wait(lk, [j] { return j->done; });
    }
    *result = j->result;
    if (queued_ns) *queued_ns = j->start_ns - j->submit_ns;
    if (run_ns) *run_ns = j->done_ns - j->start_ns;
    const int rc = j->rc;
    if (rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = j->error;
    }
    delete j;
    return rc;
}

uint32_t hj_server_slots(const hj_server* s) { return s ? (uint32_t)s->ctx.size() : 0; }

int hj_server_set_option(hj_server* s, const char* name, int64_t value) {
    if (!s || !name || value < 0) return HJ_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (std::strcmp(name, "budget_bytes") == 0) {
        s->budget = (uint64_t)value;
        s->cv_work.notify_all();
        return HJ_OK;
    }
    if (std::strcmp(name, "dir_cache_bytes") == 0) {
        std::lock_guard<std::mutex> ck(s->dcache->mu);
        s->dcache->budget = (uint64_t)value;  // 0: nothing is cached from now on
        return HJ_OK;
    }
    return HJ_E_INVALID;
}

int hj_server_cache_stats(hj_server* s, uint64_t* hits, uint64_t* misses, uint64_t* entries,
                          uint64_t* bytes) {
    if (!s || s->ctx.empty()) return HJ_E_INVALID;
    return hj_dir_cache_stats(s->ctx[0], hits, misses, entries, bytes);
}

const char* hj_server_last_error(hj_server* s) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    static thread_local std::string copy;
    copy = s->last_error;
    return copy.c_str();
}

}  // extern "C"

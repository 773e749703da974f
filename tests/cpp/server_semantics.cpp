// Request-server semantics on the CPU, with small test services: the error
// strings and behaviours the reference's server tests assert
// (test/server/Server_Test.cpp:31-270), the item-processor drain contract with
// a service that keeps items in flight across calls, and concurrent clients.
// Prints one line per check; exit status 0 when all pass.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "as/Server.h"
#include "InflightRing.h"

using Types = GenericTypes<int, int, int>;

static int failures = 0;

#define CHECK(cond, name)                                                         \
    do {                                                                          \
        if (cond) std::printf("ok   %s\n", name);                                 \
        else { std::printf("FAIL %s (line %d)\n", name, __LINE__); ++failures; } \
    } while (0)

template <class F>
static std::string error_of(F&& f) {
    try {
        f();
    } catch (std::exception const& e) {
        return e.what();
    }
    return "";
}

// result = input * common, each item finished in the call that receives it
class MultiplyService : public as::Service<Types> {
public:
    MultiplyService()
        : as::Service<Types>(std::make_shared<as::HostAllocator<int>>(),
                             std::make_shared<as::HostAllocator<int>>()) {}
    std::atomic<int> processors{0};
    itemProcessor_t createItemProcessor() override {
        ++processors;
        return [](workItem_t* item) {
            if (item) {
                for (unsigned i = 0; i < item->size(); ++i)
                    item->resultBuffer()[i] = item->inputBuffer()[i] * *item->common();
                item->setProcessed();
            }
            return false;
        };
    }
    distributor_t createDistributor() override {
        return [](int const*, size_t n) {
            std::vector<as::workerId_t> ids(n);
            std::iota(ids.begin(), ids.end(), 0u);
            return ids;
        };
    }
};

// Keeps up to three items in flight and finishes the oldest one per call, so
// completion depends on the server calling again with nullptr.
class LaggingService : public MultiplyService {
public:
    itemProcessor_t createItemProcessor() override {
        auto held = std::make_shared<std::deque<workItem_t*>>();
        return [held](workItem_t* item) {
            if (item) held->push_back(item);
            if (held->size() > 3 || (!item && !held->empty())) {
                workItem_t* it = held->front();
                held->pop_front();
                for (unsigned i = 0; i < it->size(); ++i)
                    it->resultBuffer()[i] = it->inputBuffer()[i] * *it->common();
                it->setProcessed();
            }
            return !held->empty();
        };
    }
};

// Never finishes an item (the reference's Service_TimeOut).
class StuckService : public MultiplyService {
public:
    itemProcessor_t createItemProcessor() override {
        return [](workItem_t*) { return false; };
    }
};

class NoAllocatorService : public MultiplyService {
public:
    NoAllocatorService() { setInputAllocator(nullptr); }
};

template <class S>
static bool run_sizes(const char* name, std::vector<unsigned> workers) {
    auto svc = std::make_shared<S>();
    as::Server<Types> server(svc);
    std::vector<int> input(100000), result(100000);
    std::iota(input.begin(), input.end(), 0);
    bool ok = true;
    for (unsigned w : workers) {
        server.createWorkers(w);
        for (unsigned n : {10u, 537u, 1001u, 1u, 2u, 3u, 100000u})
            for (unsigned item : {1u, 2u, 3u, 5u, 7u, 55u, 111u, 390u, 557u, 1000u}) {
                if (n == 100000u && item < 55u) continue;  // keeps the run short
                as::Request<int, int> req(input.data(), n, 3);
                server.setItemSize(item);
                server.submit(req);
                std::fill(result.begin(), result.begin() + n, -1);
                server.wait(req, result.data());
                for (unsigned i = 0; i < n; ++i) ok &= result[i] == 3 * input[i];
            }
    }
    CHECK(ok, name);
    return ok;
}

// Keeps up to three items in flight like the MI355 pipeline; every item whose
// first input is negative "fails" at completion (its slot is freed, the item
// left unprocessed).  Completion is decided through retire_in_order.
class FailingService : public MultiplyService {
public:
    itemProcessor_t createItemProcessor() override {
        struct Ring {
            std::deque<int> inflight;
            workItem_t* slots[3] = {nullptr, nullptr, nullptr};
        };
        auto r = std::make_shared<Ring>();
        return [r](workItem_t* item) {
            auto retire = [&](bool block) {
                as::retire_in_order(
                    r->inflight, block,
                    [&](int k, bool) {
                        return r->slots[k]->inputBuffer()[0] < 0 ? as::SlotStatus::Failed : as::SlotStatus::Done;
                    },
                    [&](int k) {
                        workItem_t* it = r->slots[k];
                        for (unsigned i = 0; i < it->size(); ++i)
                            it->resultBuffer()[i] = it->inputBuffer()[i] * *it->common();
                        it->setProcessed();
                        r->slots[k] = nullptr;
                    },
                    [&](int k) { r->slots[k] = nullptr; });
            };
            if (item) {
                if (r->inflight.size() == 3) retire(true);
                int k = 0;
                while (r->slots[k]) ++k;
                r->slots[k] = item;
                r->inflight.push_back(k);
            } else {
                retire(true);
            }
            return !r->inflight.empty();
        };
    }
};

int main() {
    using server_t = as::Server<Types>;
    using request_t = server_t::request_t;

    CHECK(error_of([] { server_t s(nullptr); }) == "Invalid service (nullptr).", "construct with nullptr service");
    CHECK(error_of([] { server_t s(std::make_shared<NoAllocatorService>()); }) ==
              "Invalid service allocators (nullptr).",
          "construct with missing allocators");
    {
        server_t s(std::make_shared<MultiplyService>());
        CHECK(error_of([&] { s.setService(nullptr); }) == "Invalid service (nullptr).", "setService(nullptr)");
        CHECK(error_of([&] { s.setService(std::make_shared<NoAllocatorService>()); }) ==
                  "Invalid service allocators (nullptr).",
              "setService without allocators");
        CHECK(error_of([&] { s.setItemSize(0); }) == "Item size must be greater than zero.", "setItemSize(0)");
        s.setItemSize(1368);
        CHECK(s.itemSize() == 1368, "setItemSize");
    }
    {
        auto svc = std::make_shared<MultiplyService>();
        server_t s(svc);
        s.createWorkers(1);
        request_t req;
        CHECK(error_of([&] { s.submit(req); }) == "Invalid input buffer (nullptr).", "submit without buffer");
        std::vector<int> in(1000, 2), out(1000, 0);
        req.setDataAndSize(in.data(), in.size());
        req.setCommon(5);
        s.submit(req);
        CHECK(req.data() == nullptr && req.size() == 0, "submit resets the request");
        req.setDataAndSize(in.data(), in.size());
        CHECK(error_of([&] { s.submit(req); }) == "Request has already been submitted.", "double submit");
        s.wait(req, out.data());
        CHECK(out[0] == 10 && out[999] == 10, "results after double-submit error");
        request_t other;
        CHECK(error_of([&] { s.wait(other, out.data()); }) ==
                  "Either request has not been submitted or clientBuffer invalid (nullptr).",
              "wait on a request never submitted");
        req.setDataAndSize(in.data(), 1);
        s.submit(req);
        CHECK(error_of([&] { s.wait(req, nullptr); }) ==
                  "Either request has not been submitted or clientBuffer invalid (nullptr).",
              "wait with nullptr result buffer");
        s.wait(req, out.data());
        bool ok = true;
        for (int i = 0; i < 10; ++i) {
            req.setDataAndSize(in.data(), 0);
            s.submit(req);
            s.wait(req, out.data());
        }
        CHECK(ok, "empty requests");
    }
    {
        server_t s(std::make_shared<StuckService>());
        s.createWorkers(1);
        s.setWaitTime(10);
        std::vector<int> in(1, 1), out(1);
        request_t req(in.data(), 1, 1);
        s.submit(req);
        CHECK(error_of([&] { s.wait(req, out.data()); }) == "Request takes too long to finish for unknown reasons",
              "time-out");
    }
    run_sizes<MultiplyService>("item sizes x request sizes x workers", {1, 2, 3, 4});
    run_sizes<LaggingService>("in-flight items drained with nullptr calls", {1, 3});
    {
        // several clients at once on one server
        auto svc = std::make_shared<LaggingService>();
        server_t s(svc);
        s.createWorkers(3);
        s.setItemSize(37);
        std::atomic<bool> ok{true};
        std::vector<std::thread> clients;
        for (int c = 0; c < 6; ++c)
            clients.emplace_back([&, c] {
                std::vector<int> in(5000 + 17 * c), out(in.size());
                std::iota(in.begin(), in.end(), c);
                for (int round = 0; round < 20; ++round) {
                    request_t req(in.data(), in.size(), c + 1);
                    s.submit(req);
                    s.wait(req, out.data());
                    for (size_t i = 0; i < in.size(); ++i)
                        if (out[i] != in[i] * (c + 1)) ok = false;
                }
            });
        for (auto& t : clients) t.join();
        CHECK(ok.load(), "concurrent clients");
    }
    {
        // submit several requests before waiting, wait out of order
        server_t s(std::make_shared<MultiplyService>());
        s.createWorkers(2);
        s.setItemSize(100);
        std::vector<std::vector<int>> in(5), out(5);
        std::vector<request_t> reqs(5);
        for (int r = 0; r < 5; ++r) {
            in[r].assign(1000 + r, r + 1);
            out[r].assign(in[r].size(), 0);
            reqs[r] = request_t(in[r].data(), in[r].size(), 2);
            s.submit(reqs[r]);
        }
        bool ok = true;
        for (int r = 4; r >= 0; --r) {
            s.wait(reqs[r], out[r].data());
            for (int v : out[r]) ok &= v == 2 * (r + 1);
        }
        CHECK(ok, "several requests in flight, waited out of order");
    }
    {
        // A failed completion frees its slot: the failed request times out in
        // wait(), the others complete, and the server shuts down.
        server_t s(std::make_shared<FailingService>());
        s.setItemSize(10);
        s.setWaitTime(200);
        s.createWorkers(2);
        std::vector<int> good(100, 3), bad(100, 3), outg(100, 0), outb(100, 0);
        bad[50] = -1;  // the item holding inputs 50..59 fails
        request_t rg, rb;
        rg.setDataAndSize(good.data(), good.size());
        rg.setCommon(2);
        rb.setDataAndSize(bad.data(), bad.size());
        rb.setCommon(2);
        s.submit(rg);
        s.submit(rb);
        s.wait(rg, outg.data());
        CHECK(outg[0] == 6 && outg[99] == 6, "failed slots: other requests complete");
        CHECK(error_of([&] { s.wait(rb, outb.data()); }) ==
                  "Request takes too long to finish for unknown reasons",
              "failed slots: the failed request times out");
    }
    CHECK(true, "failed slots: server shut down");
    std::printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}

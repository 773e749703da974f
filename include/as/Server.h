/*
 * Request server with the reference's interface and error behaviour
 * (server/Server.h:42-103, server/Server.tpp:43-206): submit() splits a
 * request into work items of itemSize() inputs and hands them to the workers
 * the service's distributor names; wait() returns the results in input order.
 *
 * The machinery is this library's own: submit() dispatches synchronously
 * (each item to the eligible worker with the shortest queue), every worker
 * owns a queue and the item processor the service created for it, and wait()
 * sleeps on a progress signal the workers raise after every processor call
 * instead of polling, bounded by setWaitTime() milliseconds.  Server buffers
 * come from the service's allocators and are pooled by capacity.
 *
 * Thread safety: submit() and wait() may be called from several client
 * threads; one Request object belongs to one client at a time.
 */
#ifndef AS_SERVER_H_
#define AS_SERVER_H_

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <vector>

#include "Service.h"

namespace as {

constexpr unsigned DEFAULT_ITEM_SIZE = 100;       // server/Server.h:39
constexpr unsigned MAX_WAIT_MILLISECONDS = 20000; // server/Server.h:40

template <typename GenericTypes>
class Server {
public:
    using service_t = Service<GenericTypes>;
    using input_t = typename GenericTypes::input_t;
    using result_t = typename GenericTypes::result_t;
    using common_t = typename GenericTypes::common_t;
    using workItem_t = typename service_t::workItem_t;
    using request_t = Request<input_t, common_t>;

    explicit Server(std::shared_ptr<service_t> service) { setService(service); }
    Server(Server const&) = delete;
    Server& operator=(Server const&) = delete;

    ~Server() {
        stopWorkers();
        std::lock_guard<std::mutex> g(_mutex);
        for (auto& kv : _pending) releaseBuffers(*kv.second);
        for (auto& p : _abandoned) releaseBuffers(*p);
        for (auto& kv : _inPool) _inAlloc->deallocate(kv.second, kv.first);
        for (auto& kv : _outPool) _outAlloc->deallocate(kv.second, kv.first);
    }

    std::shared_ptr<service_t> service() noexcept { return _service; }

    void setService(std::shared_ptr<service_t> const& service) {
        if (service.get() == nullptr) throw std::invalid_argument("Invalid service (nullptr).");
        if (service->getInputAllocator() == nullptr || service->getResultAllocator() == nullptr)
            throw std::invalid_argument("Invalid service allocators (nullptr).");
        std::lock_guard<std::mutex> g(_mutex);
        if (!_inPool.empty() || !_outPool.empty()) {
            for (auto& kv : _inPool) _inAlloc->deallocate(kv.second, kv.first);
            for (auto& kv : _outPool) _outAlloc->deallocate(kv.second, kv.first);
            _inPool.clear();
            _outPool.clear();
        }
        _service = service;
        _inAlloc = service->getInputAllocator();
        _outAlloc = service->getResultAllocator();
        _distributor = service->createDistributor();
    }

    /* Replaces the worker pool; call it when no request is in flight. */
    void createWorkers(unsigned number) {
        stopWorkers();
        std::vector<std::unique_ptr<Worker>> pool;
        for (unsigned i = 0; i < number; ++i) {
            std::unique_ptr<Worker> w(new Worker);
            w->process = _service->createItemProcessor();
            pool.push_back(std::move(w));
        }
        for (auto& w : pool) {
            Worker* wp = w.get();
            wp->thread = std::thread([this, wp] { run(*wp); });
        }
        std::lock_guard<std::mutex> g(_mutex);
        _workers = std::move(pool);
    }

    void setWaitTime(unsigned milliSeconds) noexcept { _waitTime = milliSeconds; }

    void setItemSize(size_t itemSize) {
        if (itemSize == 0) throw std::invalid_argument("Item size must be greater than zero.");
        _itemSize = itemSize;
    }
    size_t itemSize() const noexcept { return _itemSize; }

    void submit(request_t& request) {
        std::unique_ptr<Pending> p(new Pending);
        {
            std::lock_guard<std::mutex> g(_mutex);
            if (_pending.count(&request))
                throw std::invalid_argument("Request has already been submitted.");
            if (request.data() == nullptr)
                throw std::invalid_argument("Invalid input buffer (nullptr).");
            p->size = request.size();
            p->common = request.common();
            if (p->size > 0) {
                if (_workers.empty())
                    throw std::runtime_error("No workers (createWorkers was not called).");
                p->in = take(_inPool, _inAlloc, p->size, p->inCap);
                p->out = take(_outPool, _outAlloc, p->size, p->outCap);
            }
        }
        if (p->size > 0) {
            std::copy(request.data(), request.data() + p->size, p->in);
            const size_t nItems = (p->size + _itemSize - 1) / _itemSize;
            p->items.reset(new workItem_t[nItems]);
            p->numItems = nItems;
            for (size_t k = 0; k < nItems; ++k) {
                const size_t lo = k * _itemSize, sz = std::min(_itemSize, p->size - lo);
                p->items[k].setDataAndSize(p->in + lo, (unsigned)sz);
                p->items[k].setResultPtr(p->out + lo);
                p->items[k].setCommonPtr(&p->common);
            }
        }
        Pending* raw = p.get();
        {
            std::lock_guard<std::mutex> g(_mutex);
            _pending[&request] = std::move(p);
        }
        if (raw->numItems) dispatch(*raw);
        request.reset();
    }

    void wait(request_t const& request, result_t* result) {
        Pending* p = nullptr;
        {
            std::lock_guard<std::mutex> g(_mutex);
            auto it = _pending.find(&request);
            if (it != _pending.end()) p = it->second.get();
        }
        if (p == nullptr || result == nullptr)
            throw std::invalid_argument(
                "Either request has not been submitted or clientBuffer invalid (nullptr).");
        if (!waitProcessed(*p)) {
            // The workers may still hold the items: keep the buffers until the
            // server ends, forget the request.
            std::lock_guard<std::mutex> g(_mutex);
            auto it = _pending.find(&request);
            _abandoned.push_back(std::move(it->second));
            _pending.erase(it);
            throw std::runtime_error("Request takes too long to finish for unknown reasons");
        }
        if (p->size > 0) std::copy(p->out, p->out + p->size, result);
        std::lock_guard<std::mutex> g(_mutex);
        auto it = _pending.find(&request);
        releaseBuffers(*it->second);
        _pending.erase(it);
    }

private:
    struct Pending {
        size_t size = 0;
        common_t common{};
        input_t* in = nullptr;
        result_t* out = nullptr;
        size_t inCap = 0, outCap = 0;
        std::unique_ptr<workItem_t[]> items;
        size_t numItems = 0;
        size_t firstOpen = 0;  // items before this index are known to be processed
    };

    struct Worker {
        std::thread thread;
        std::mutex m;
        std::condition_variable cv;
        std::deque<workItem_t*> queue;
        bool stop = false;
        typename service_t::itemProcessor_t process;
    };

    template <class T>
    static T* take(std::multimap<size_t, T*>& pool, Allocator<T>* alloc, size_t n, size_t& cap) {
        auto it = pool.lower_bound(n);
        if (it != pool.end()) {
            cap = it->first;
            T* p = it->second;
            pool.erase(it);
            return p;
        }
        cap = n;
        return alloc->allocate(n);
    }

    void releaseBuffers(Pending& p) {
        if (p.in) _inPool.emplace(p.inCap, p.in);
        if (p.out) _outPool.emplace(p.outCap, p.out);
        p.in = nullptr;
        p.out = nullptr;
    }

    void dispatch(Pending& p) {
        std::vector<Worker*> eligible;
        {
            std::lock_guard<std::mutex> g(_mutex);
            const std::vector<workerId_t> ids = _distributor(&p.common, _workers.size());
            for (workerId_t id : ids)
                if (id < _workers.size()) eligible.push_back(_workers[id].get());
            if (eligible.empty()) {
                for (auto& w : _workers) eligible.push_back(w.get());
            }
        }
        std::vector<size_t> load(eligible.size());
        for (size_t k = 0; k < eligible.size(); ++k) {
            std::lock_guard<std::mutex> g(eligible[k]->m);
            load[k] = eligible[k]->queue.size();
        }
        std::vector<std::vector<workItem_t*>> batch(eligible.size());
        for (size_t i = 0; i < p.numItems; ++i) {
            const size_t k = std::min_element(load.begin(), load.end()) - load.begin();
            batch[k].push_back(&p.items[i]);
            ++load[k];
        }
        for (size_t k = 0; k < eligible.size(); ++k) {
            if (batch[k].empty()) continue;
            {
                std::lock_guard<std::mutex> g(eligible[k]->m);
                eligible[k]->queue.insert(eligible[k]->queue.end(), batch[k].begin(), batch[k].end());
            }
            eligible[k]->cv.notify_one();
        }
    }

    bool allProcessed(Pending& p) {
        while (p.firstOpen < p.numItems && p.items[p.firstOpen].isProcessed()) ++p.firstOpen;
        return p.firstOpen == p.numItems;
    }

    bool waitProcessed(Pending& p) {
        const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(_waitTime);
        std::unique_lock<std::mutex> lk(_progressMutex);
        while (!allProcessed(p)) {
            const auto now = std::chrono::steady_clock::now();
            if (now >= deadline) return allProcessed(p);
            // bounded sleep: a processor may also finish items from a thread
            // that does not raise the progress signal
            _progress.wait_until(lk, std::min(deadline, now + std::chrono::milliseconds(2)));
        }
        return true;
    }

    void run(Worker& w) {
        bool again = false;
        for (;;) {
            workItem_t* item = nullptr;
            {
                std::unique_lock<std::mutex> lk(w.m);
                if (!again) {
                    w.cv.wait(lk, [&] { return w.stop || !w.queue.empty(); });
                    if (w.queue.empty()) break;  // stop requested, nothing in flight
                }
                if (!w.queue.empty()) {
                    item = w.queue.front();
                    w.queue.pop_front();
                }
            }
            again = w.process(item);
            {
                std::lock_guard<std::mutex> g(_progressMutex);
            }
            _progress.notify_all();
        }
    }

    void stopWorkers() {
        std::vector<std::unique_ptr<Worker>> pool;
        {
            std::lock_guard<std::mutex> g(_mutex);
            pool = std::move(_workers);
            _workers.clear();
        }
        for (auto& w : pool) {
            {
                std::lock_guard<std::mutex> g(w->m);
                w->stop = true;
            }
            w->cv.notify_all();
        }
        for (auto& w : pool)
            if (w->thread.joinable()) w->thread.join();
    }

    std::shared_ptr<service_t> _service;
    Allocator<input_t>* _inAlloc = nullptr;
    Allocator<result_t>* _outAlloc = nullptr;
    typename service_t::distributor_t _distributor;
    size_t _itemSize = DEFAULT_ITEM_SIZE;
    unsigned _waitTime = MAX_WAIT_MILLISECONDS;

    std::mutex _mutex;  // _pending, pools, _workers
    std::unordered_map<const request_t*, std::unique_ptr<Pending>> _pending;
    std::vector<std::unique_ptr<Pending>> _abandoned;
    std::multimap<size_t, input_t*> _inPool;
    std::multimap<size_t, result_t*> _outPool;
    std::vector<std::unique_ptr<Worker>> _workers;

    std::mutex _progressMutex;
    std::condition_variable _progress;
};

}  // namespace as

#endif

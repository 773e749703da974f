/*
 * The plugin contract of the request server: Request, WorkItem, Allocator and
 * Service keep the reference's names and meaning (server/Request.h:15-57,
 * server/WorkItem.h:16-100, allocator/Allocator.h:17-48,
 * service/Service.h:24-70), so a service written for one runs on the other.
 *
 * - A Request is non-owning: (data, size, common).  Server::submit copies the
 *   data into server-owned buffers from the service's allocators and resets
 *   the request.
 * - The item processor is called on one worker thread, with a new item or with
 *   nullptr when its previous call returned true (items still in flight) and
 *   no new item is queued.  For every finished item it writes size() results
 *   to resultBuffer() and calls setProcessed().  It returns true while items
 *   are in flight.
 * - The distributor maps (common, number of workers) to the ids of the
 *   workers allowed to take the request's items.
 */
#ifndef AS_SERVICE_H_
#define AS_SERVICE_H_

#include <atomic>
#include <cstdlib>
#include <functional>
#include <memory>
#include <new>
#include <vector>

#include "Types_6D.h"

namespace as {

template <typename BufferType>
class Allocator {
public:
    virtual ~Allocator() {}
    virtual BufferType* allocate(size_t count) = 0;
    virtual void deallocate(BufferType* p, size_t count = 0) = 0;
};

/* Plain host memory. */
template <typename BufferType>
class HostAllocator : public Allocator<BufferType> {
public:
    BufferType* allocate(size_t count) override {
        void* p = std::malloc(count * sizeof(BufferType));
        if (!p && count) throw std::bad_alloc();
        return static_cast<BufferType*>(p);
    }
    void deallocate(BufferType* p, size_t = 0) override { std::free(p); }
};

template <typename InputType, typename CommonType>
class Request {
public:
    Request() = default;
    Request(InputType* data, size_t size, CommonType common)
        : _data(data), _size(size), _common(common) {}

    size_t size() const noexcept { return _size; }
    InputType* data() const noexcept { return _data; }
    CommonType common() const noexcept { return _common; }

    void setDataAndSize(InputType* data, size_t size) noexcept {
        _data = data;
        _size = size;
    }
    void setCommon(CommonType const& common) noexcept { _common = common; }
    void reset() {
        _data = nullptr;
        _size = 0;
        _common = CommonType();
    }

private:
    InputType* _data = nullptr;
    size_t _size = 0;
    CommonType _common = CommonType();
};

template <typename InputType, typename CommonType, typename ResultType>
class WorkItem {
public:
    WorkItem() = default;
    WorkItem(InputType* data, unsigned size, CommonType* common, ResultType* result)
        : _data(data), _size(size), _common(common), _result(result) {}
    WorkItem(WorkItem const&) = delete;
    WorkItem& operator=(WorkItem const&) = delete;

    unsigned size() const noexcept { return _size; }
    InputType* inputBuffer() const noexcept { return _data; }
    CommonType* common() const noexcept { return _common; }
    ResultType* resultBuffer() const noexcept { return _result; }

    /* Results are visible to any thread that then observes isProcessed(). */
    void setProcessed() noexcept { _processed.store(true, std::memory_order_release); }
    bool isProcessed() const noexcept { return _processed.load(std::memory_order_acquire); }

    void setDataAndSize(InputType* data, unsigned size) noexcept {
        _data = data;
        _size = size;
    }
    void setCommonPtr(CommonType* common) noexcept { _common = common; }
    void setResultPtr(ResultType* result) noexcept { _result = result; }
    bool isValid() const noexcept {
        return _data != nullptr && _size > 0 && _common != nullptr && _result != nullptr;
    }

private:
    InputType* _data = nullptr;
    unsigned _size = 0;
    CommonType* _common = nullptr;
    ResultType* _result = nullptr;
    std::atomic<bool> _processed{false};
};

template <typename GenericTypes>
class Service {
public:
    using genericTypes_t = GenericTypes;
    using service_t = Service<GenericTypes>;
    using input_t = typename GenericTypes::input_t;
    using common_t = typename GenericTypes::common_t;
    using result_t = typename GenericTypes::result_t;
    using workItem_t = WorkItem<input_t, common_t, result_t>;
    using itemProcessor_t = std::function<bool(workItem_t*)>;
    using distributor_t = std::function<std::vector<workerId_t>(common_t const*, size_t)>;

    Service(std::shared_ptr<Allocator<input_t>> inputAllocator,
            std::shared_ptr<Allocator<result_t>> resultAllocator)
        : _inputAllocator(std::move(inputAllocator)), _resultAllocator(std::move(resultAllocator)) {}
    virtual ~Service() {}

    void setInputAllocator(std::shared_ptr<Allocator<input_t>> a) { _inputAllocator = std::move(a); }
    void setResultAllocator(std::shared_ptr<Allocator<result_t>> a) { _resultAllocator = std::move(a); }
    Allocator<input_t>* getInputAllocator() { return _inputAllocator.get(); }
    Allocator<result_t>* getResultAllocator() { return _resultAllocator.get(); }

    virtual itemProcessor_t createItemProcessor() = 0;
    virtual distributor_t createDistributor() = 0;

private:
    std::shared_ptr<Allocator<input_t>> _inputAllocator;
    std::shared_ptr<Allocator<result_t>> _resultAllocator;
};

}  // namespace as

#endif

// Scoring services over the C-ABI: per-device contexts with resident data,
// and a per-worker item pipeline of kSlots streams (the role of
// GPUEnergyService6D::Private's stage pipeline, GPUEnergyService6D.tpp:94-457,
// with one fused launch per item instead of five stages).
#include "as/MI355EnergyService6D.h"

#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>

#include "InflightRing.h"

namespace as {

struct DeviceSet {
    struct Dev {
        int hipDevice = 0;
        att_ctx* ctx = nullptr;
        std::mutex m;  // every call on ctx; the maps below
        std::map<id_t, att_handle> proteins, grids;
        std::map<std::pair<id_t, id_t>, att_handle> params;
    };
    std::vector<std::unique_ptr<Dev>> devs;

    explicit DeviceSet(std::vector<int> const& ids) {
        if (ids.empty()) throw std::invalid_argument("No device ids given.");
        for (int id : ids) {
            std::unique_ptr<Dev> d(new Dev);
            d->hipDevice = id;
            if (att_ctx_create(id, &d->ctx) != ATT_OK)
                throw std::runtime_error("Cannot create a context on HIP device " + std::to_string(id) + ".");
            devs.push_back(std::move(d));
        }
    }
    ~DeviceSet() {
        for (auto& d : devs) att_ctx_destroy(d->ctx);  // releases every upload
    }
};

namespace {

constexpr int kSlots = 4;

void check(att_ctx* ctx, att_status st, const char* what) {
    if (st != ATT_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? att_last_error(ctx) : "no context"));
}

template <class T>
std::shared_ptr<T> item_as(DataManager& dm, id_t id, const char* kind) {
    auto p = std::dynamic_pointer_cast<T>(dm.get(id));
    if (!p)
        throw std::invalid_argument("DataItem (id = " + std::to_string(id) + ") is not a " + kind + ".");
    return p;
}

// Upload helpers; the caller holds d.m.
att_handle protein_handle(DeviceSet::Dev& d, DataManager& dm, id_t id) {
    auto it = d.proteins.find(id);
    if (it != d.proteins.end()) return it->second;
    auto p = item_as<ProteinData>(dm, id, "protein");
    const att_protein_desc desc = p->desc();
    att_handle h;
    check(d.ctx, att_upload_protein(d.ctx, &desc, &h), "upload protein");
    d.proteins[id] = h;
    return h;
}

att_handle grid_handle(DeviceSet::Dev& d, DataManager& dm, id_t id) {
    auto it = d.grids.find(id);
    if (it != d.grids.end()) return it->second;
    auto g = item_as<GridUnionData>(dm, id, "grid union");
    const att_grid_desc in = g->innerDesc(), out = g->outerDesc();
    const att_nlgrid_desc nl = g->nlDesc();
    att_handle h;
    check(d.ctx, att_upload_grid(d.ctx, &in, &out, g->hasNL ? &nl : nullptr, &h), "upload grid");
    d.grids[id] = h;
    return h;
}

att_handle params_handle(DeviceSet::Dev& d, DataManager& dm, id_t tableId, id_t paramsId) {
    auto key = std::make_pair(tableId, paramsId);
    auto it = d.params.find(key);
    if (it != d.params.end()) return it->second;
    auto t = item_as<ParamTableData>(dm, tableId, "parameter table");
    auto s = item_as<SimParamData>(dm, paramsId, "simulation parameter set");
    att_handle h;
    check(d.ctx, att_upload_params(d.ctx, t->table.data(), t->numTypes, t->potShape, s->dielec,
                                   s->epsilon, s->ffelec, &h),
          "upload parameters");
    d.params[key] = h;
    return h;
}

struct RigidKind {
    using types = Types_6D<float>;
    static constexpr const char* name = "MI355EnergyService6D";
    static att_status launch(DeviceSet::Dev& d, DataManager& dm, Common const& c, const void* in,
                             size_t n, void* out, void* stream) {
        const att_handle rec = protein_handle(d, dm, c.recId), lig = protein_handle(d, dm, c.ligId);
        const att_handle grid = grid_handle(d, dm, c.gridId);
        const att_handle par = params_handle(d, dm, c.tableId, c.paramsId);
        return att_score_rigid_dev(d.ctx, rec, lig, grid, par, (const att_dof6*)in, n,
                                   (att_result6*)out, stream);
    }
    static void attach(DeviceSet::Dev& d, DataManager& dm, Common const& c) {
        protein_handle(d, dm, c.recId);
        protein_handle(d, dm, c.ligId);
        grid_handle(d, dm, c.gridId);
        params_handle(d, dm, c.tableId, c.paramsId);
    }
};

struct ModesKind {
    using types = Types_6D_Modes<float>;
    static constexpr const char* name = "MI355EnergyService6DModes";
    static att_status launch(DeviceSet::Dev& d, DataManager& dm, Common_Modes const& c,
                             const void* in, size_t n, void* out, void* stream) {
        const att_handle rec = protein_handle(d, dm, c.recId), lig = protein_handle(d, dm, c.ligId);
        const att_handle rg = grid_handle(d, dm, c.gridIdRec), lg = grid_handle(d, dm, c.gridIdLig);
        const att_handle par = params_handle(d, dm, c.tableId, c.paramsId);
        return att_score_modes_dev(d.ctx, rec, lig, rg, lg, par, (const att_dof6m*)in, n,
                                   (att_result6m*)out, stream);
    }
    static void attach(DeviceSet::Dev& d, DataManager& dm, Common_Modes const& c) {
        protein_handle(d, dm, c.recId);
        protein_handle(d, dm, c.ligId);
        grid_handle(d, dm, c.gridIdRec);
        grid_handle(d, dm, c.gridIdLig);
        params_handle(d, dm, c.tableId, c.paramsId);
    }
};

// Items in flight on one worker: each slot owns a stream, an event and
// device buffers for one item; slots retire in issue order.
template <class Kind>
class Pipeline {
    using input_t = typename Kind::types::input_t;
    using result_t = typename Kind::types::result_t;
    using common_t = typename Kind::types::common_t;
    using item_t = WorkItem<input_t, common_t, result_t>;

    struct Slot {
        void* stream = nullptr;
        void* done = nullptr;
        void* d_in = nullptr;
        void* d_out = nullptr;
        size_t cap = 0;
        item_t* item = nullptr;
    };

public:
    Pipeline(std::shared_ptr<DeviceSet> devs, DeviceSet::Dev& dev, std::shared_ptr<DataManager> dm)
        : _devs(std::move(devs)), _dev(dev), _dm(std::move(dm)) {
        std::lock_guard<std::mutex> g(_dev.m);
        for (auto& s : _slots) {
            check(_dev.ctx, att_stream_create(_dev.ctx, &s.stream), "stream");
            check(_dev.ctx, att_event_create(_dev.ctx, &s.done), "event");
        }
    }
    ~Pipeline() {
        std::lock_guard<std::mutex> g(_dev.m);
        for (auto& s : _slots) {
            if (s.stream) att_sync(_dev.ctx, s.stream);
            if (s.d_in) att_dev_free(_dev.ctx, s.d_in);
            if (s.d_out) att_dev_free(_dev.ctx, s.d_out);
            if (s.done) att_event_destroy(_dev.ctx, s.done);
            if (s.stream) att_stream_destroy(_dev.ctx, s.stream);
        }
    }

    bool operator()(item_t* item) {
        try {
            retire(false);
            if (item != nullptr) {
                if ((int)_inflight.size() == kSlots) retire(true);
                issue(item);
            } else if (!_inflight.empty()) {
                retire(true);
            }
        } catch (std::exception const& e) {
            // The item-processor contract has no error channel: the failed
            // item stays unprocessed and the client's wait() reports it.
            std::fprintf(stderr, "%s: %s\n", Kind::name, e.what());
        }
        return !_inflight.empty();
    }

private:
    void issue(item_t* item) {
        int k = 0;
        while (_slots[k].item != nullptr) ++k;  // a free slot exists: fewer than kSlots in flight
        Slot& s = _slots[k];
        const size_t n = item->size();
        std::lock_guard<std::mutex> g(_dev.m);
        if (s.cap < n) {
            if (s.d_in) att_dev_free(_dev.ctx, s.d_in);
            if (s.d_out) att_dev_free(_dev.ctx, s.d_out);
            s.d_in = s.d_out = nullptr;
            s.cap = 0;
            check(_dev.ctx, att_dev_alloc(_dev.ctx, n * sizeof(input_t), &s.d_in), "device buffer");
            check(_dev.ctx, att_dev_alloc(_dev.ctx, n * sizeof(result_t), &s.d_out), "device buffer");
            s.cap = n;
        }
        check(_dev.ctx, att_memcpy_h2d_async(_dev.ctx, s.d_in, item->inputBuffer(), n * sizeof(input_t), s.stream),
              "input copy");
        check(_dev.ctx, Kind::launch(_dev, *_dm, *item->common(), s.d_in, n, s.d_out, s.stream), "scoring");
        check(_dev.ctx, att_memcpy_d2h_async(_dev.ctx, item->resultBuffer(), s.d_out, n * sizeof(result_t), s.stream),
              "result copy");
        check(_dev.ctx, att_event_record(_dev.ctx, s.done, s.stream), "event");
        s.item = item;
        _inflight.push_back(k);
    }

    // Marks finished items processed, in issue order; with block = true waits
    // for the oldest one first.  A slot whose completion reports an error is
    // freed with its item left unprocessed, then the error is raised.
    void retire(bool block) {
        att_status failed = ATT_OK;
        const bool ok = retire_in_order(
            _inflight, block,
            [&](int k, bool b) {
                // event queries and waits are thread-safe in the runtime: no
                // lock, so a blocked worker does not hold up the others
                const att_status st = b ? att_event_sync(_dev.ctx, _slots[k].done)
                                        : att_event_query(_dev.ctx, _slots[k].done);
                if (st == ATT_OK) return SlotStatus::Done;
                if (st == ATT_NOT_READY) return SlotStatus::Pending;
                failed = st;
                return SlotStatus::Failed;
            },
            [&](int k) {
                _slots[k].item->setProcessed();
                _slots[k].item = nullptr;
            },
            [&](int k) { _slots[k].item = nullptr; });
        if (!ok) {
            std::lock_guard<std::mutex> g(_dev.m);
            check(_dev.ctx, failed, "completion");
        }
    }

    std::shared_ptr<DeviceSet> _devs;
    DeviceSet::Dev& _dev;
    std::shared_ptr<DataManager> _dm;
    Slot _slots[kSlots];
    std::deque<int> _inflight;
};

template <class Kind>
std::function<bool(WorkItem<typename Kind::types::input_t, typename Kind::types::common_t,
                            typename Kind::types::result_t>*)>
make_processor(std::shared_ptr<DeviceSet> const& devs, unsigned worker, std::shared_ptr<DataManager> const& dm) {
    DeviceSet::Dev& dev = *devs->devs[worker % devs->devs.size()];
    auto p = std::make_shared<Pipeline<Kind>>(devs, dev, dm);
    return [p](WorkItem<typename Kind::types::input_t, typename Kind::types::common_t,
                        typename Kind::types::result_t>* item) { return (*p)(item); };
}

std::vector<workerId_t> all_workers(size_t numWorkers) {
    std::vector<workerId_t> ids(numWorkers);
    std::iota(ids.begin(), ids.end(), 0u);
    return ids;
}

}  // namespace

template <typename BufferType>
BufferType* HostPinnedAllocator<BufferType>::allocate(size_t count) {
    DeviceSet::Dev& d = *_devices->devs[0];
    void* p = nullptr;
    std::lock_guard<std::mutex> g(d.m);
    check(d.ctx, att_host_alloc(d.ctx, count * sizeof(BufferType), &p), "pinned host buffer");
    return static_cast<BufferType*>(p);
}

template <typename BufferType>
void HostPinnedAllocator<BufferType>::deallocate(BufferType* p, size_t) {
    if (!p) return;
    DeviceSet::Dev& d = *_devices->devs[0];
    std::lock_guard<std::mutex> g(d.m);
    att_host_free(d.ctx, p);
}

// ---- rigid ----------------------------------------------------------------

template <typename REAL>
MI355EnergyService6D<REAL>::MI355EnergyService6D(std::shared_ptr<DataManager> dataMng,
                                                 std::vector<int> const& deviceIds)
    : service_t(nullptr, nullptr), _dataMng(std::move(dataMng)),
      _devices(std::make_shared<DeviceSet>(deviceIds)) {
    static_assert(sizeof(REAL) == 4, "the scoring kernels compute in FP32");
    this->setInputAllocator(std::make_shared<HostPinnedAllocator<input_t>>(_devices));
    this->setResultAllocator(std::make_shared<HostPinnedAllocator<result_t>>(_devices));
}

template <typename REAL>
MI355EnergyService6D<REAL>::~MI355EnergyService6D() {}

template <typename REAL>
auto MI355EnergyService6D<REAL>::createItemProcessor() -> itemProcessor_t {
    return make_processor<RigidKind>(_devices, _nextWorker++, _dataMng);
}

template <typename REAL>
auto MI355EnergyService6D<REAL>::createDistributor() -> distributor_t {
    return [](common_t const*, size_t numWorkers) { return all_workers(numWorkers); };
}

template <typename REAL>
void MI355EnergyService6D<REAL>::attach(common_t const& common, deviceId_t deviceId) {
    DeviceSet::Dev& d = *_devices->devs.at(deviceId);
    std::lock_guard<std::mutex> g(d.m);
    RigidKind::attach(d, *_dataMng, common);
}

// ---- modes ----------------------------------------------------------------

template <typename REAL>
MI355EnergyService6DModes<REAL>::MI355EnergyService6DModes(std::shared_ptr<DataManager> dataMng,
                                                           std::vector<int> const& deviceIds)
    : service_t(nullptr, nullptr), _dataMng(std::move(dataMng)),
      _devices(std::make_shared<DeviceSet>(deviceIds)) {
    static_assert(sizeof(REAL) == 4, "the scoring kernels compute in FP32");
    this->setInputAllocator(std::make_shared<HostPinnedAllocator<input_t>>(_devices));
    this->setResultAllocator(std::make_shared<HostPinnedAllocator<result_t>>(_devices));
}

template <typename REAL>
MI355EnergyService6DModes<REAL>::~MI355EnergyService6DModes() {}

template <typename REAL>
auto MI355EnergyService6DModes<REAL>::createItemProcessor() -> itemProcessor_t {
    return make_processor<ModesKind>(_devices, _nextWorker++, _dataMng);
}

template <typename REAL>
auto MI355EnergyService6DModes<REAL>::createDistributor() -> distributor_t {
    return [](common_t const*, size_t numWorkers) { return all_workers(numWorkers); };
}

template <typename REAL>
void MI355EnergyService6DModes<REAL>::attach(common_t const& common, deviceId_t deviceId) {
    DeviceSet::Dev& d = *_devices->devs.at(deviceId);
    std::lock_guard<std::mutex> g(d.m);
    ModesKind::attach(d, *_dataMng, common);
}

template class HostPinnedAllocator<DOF_6D<float>>;
template class HostPinnedAllocator<Result_6D<float>>;
template class HostPinnedAllocator<DOF_6D_Modes<float>>;
template class HostPinnedAllocator<Result_6D_Modes<float>>;
template class MI355EnergyService6D<float>;
template class MI355EnergyService6DModes<float>;

}  // namespace as

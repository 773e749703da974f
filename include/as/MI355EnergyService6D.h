/*
 * Scoring services for the request server, drop-in replacements of
 * GPUEnergyService6D<float> / GPUEnergyService6DModes<float>
 * (service/GPUEnergyService6D.h:29-75, service/GPUEnergyService6DModes.h)
 * over the C-ABI of libattract_hip.
 *
 * - Worker k runs on device deviceIds[k % deviceIds.size()] (the reference
 *   binds worker index to device, GPUEnergyService6D.tpp:470-477).
 * - Data registered in the DataManager is uploaded to a device on first use
 *   (or by attachAllDataToDevice) and stays resident until the service ends.
 * - The item processor keeps up to kSlots items in flight per worker, each on
 *   its own stream: input copy from the pinned server buffer, one fused
 *   scoring launch, result copy back into the server buffer; an item is
 *   marked processed once its result copy has completed.
 * - Server buffers are pinned host memory (HostPinnedAllocator), as the
 *   reference's GPU services use (GPUEnergyService.tpp:20-24).
 * - Library errors become std::runtime_error carrying the library's message.
 */
#ifndef AS_MI355ENERGYSERVICE6D_H_
#define AS_MI355ENERGYSERVICE6D_H_

#include <memory>
#include <vector>

#include "DataManager.h"
#include "Service.h"

namespace as {

struct DeviceSet;  // per-device contexts and resident data (service/MI355EnergyService6D.cpp)

/* Pinned host memory on the first device of a DeviceSet. */
template <typename BufferType>
class HostPinnedAllocator : public Allocator<BufferType> {
public:
    explicit HostPinnedAllocator(std::shared_ptr<DeviceSet> devices) : _devices(std::move(devices)) {}
    BufferType* allocate(size_t count) override;
    void deallocate(BufferType* p, size_t count = 0) override;

private:
    std::shared_ptr<DeviceSet> _devices;
};

template <typename REAL>
class MI355EnergyService6D : public Service<Types_6D<REAL>> {
public:
    using service_t = Service<Types_6D<REAL>>;
    using typename service_t::input_t;
    using typename service_t::common_t;
    using typename service_t::result_t;
    using typename service_t::workItem_t;
    using typename service_t::itemProcessor_t;
    using typename service_t::distributor_t;

    MI355EnergyService6D(std::shared_ptr<DataManager> dataMng, std::vector<int> const& deviceIds);
    ~MI355EnergyService6D() override;

    itemProcessor_t createItemProcessor() override;
    distributor_t createDistributor() override;

    /* Uploads every item the common refers to (optional; done on first use). */
    void attach(common_t const& common, deviceId_t deviceId);

private:
    std::shared_ptr<DataManager> _dataMng;
    std::shared_ptr<DeviceSet> _devices;
    unsigned _nextWorker = 0;
};

template <typename REAL>
class MI355EnergyService6DModes : public Service<Types_6D_Modes<REAL>> {
public:
    using service_t = Service<Types_6D_Modes<REAL>>;
    using typename service_t::input_t;
    using typename service_t::common_t;
    using typename service_t::result_t;
    using typename service_t::workItem_t;
    using typename service_t::itemProcessor_t;
    using typename service_t::distributor_t;

    MI355EnergyService6DModes(std::shared_ptr<DataManager> dataMng, std::vector<int> const& deviceIds);
    ~MI355EnergyService6DModes() override;

    itemProcessor_t createItemProcessor() override;
    distributor_t createDistributor() override;

    void attach(common_t const& common, deviceId_t deviceId);

private:
    std::shared_ptr<DataManager> _dataMng;
    std::shared_ptr<DeviceSet> _devices;
    unsigned _nextWorker = 0;
};

}  // namespace as

#endif

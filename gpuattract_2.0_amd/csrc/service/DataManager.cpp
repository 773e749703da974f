// Host-side data registry (server/DataManager.cpp:19-60 for the interface and
// error strings) and the descriptors the C-ABI uploads from.
#include "as/DataManager.h"

#include <stdexcept>
#include <string>

namespace as {

namespace {
std::invalid_argument bad_id(id_t id) {
    return std::invalid_argument("Invalid id for DataItem (id = " + std::to_string(id) + ").");
}
}  // namespace

att_protein_desc ProteinData::desc() const {
    att_protein_desc d{};
    d.n = n;
    d.numModes = numModes;
    d.xyz = xyz.data();
    d.type = type.data();
    d.mappedType = mappedType.data();
    d.charge = charge.data();
    d.modes = numModes ? modes.data() : nullptr;
    d.modeForce = numModes ? modeForce.data() : nullptr;
    return d;
}

static att_grid_desc grid_desc(const GridUnionData::Grid& g) {
    att_grid_desc d{};
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz;
    d.numGrids = g.numGrids;
    d.ox = g.ox; d.oy = g.oy; d.oz = g.oz;
    d.spacing = g.spacing;
    d.data = g.data.data();
    return d;
}

att_grid_desc GridUnionData::innerDesc() const { return grid_desc(inner); }
att_grid_desc GridUnionData::outerDesc() const { return grid_desc(outer); }

att_nlgrid_desc GridUnionData::nlDesc() const {
    att_nlgrid_desc d{};
    d.nx = nl.nx; d.ny = nl.ny; d.nz = nl.nz;
    d.ox = nl.ox; d.oy = nl.oy; d.oz = nl.oz;
    d.spacing = nl.spacing;
    d.plateau = nl.plateau;
    d.desc = nl.desc.data();
    d.list = nl.list.data();
    d.numEl = nl.list.size();
    return d;
}

id_t DataManager::add(std::shared_ptr<DataItem> item) {
    if (item.get() == nullptr) throw std::invalid_argument("Invalid DataItem (nullptr).");
    std::lock_guard<std::mutex> g(_m);
    _items.push_back(std::move(item));
    return (id_t)(_items.size() - 1);
}

std::vector<id_t> DataManager::add(std::vector<std::shared_ptr<DataItem>> const& items) {
    std::vector<id_t> ids;
    ids.reserve(items.size());
    for (auto const& it : items) ids.push_back(add(it));
    return ids;
}

bool DataManager::isValid(id_t id) const { return id < _items.size() && _items[id] != nullptr; }

std::shared_ptr<DataItem> DataManager::get(id_t id) {
    std::lock_guard<std::mutex> g(_m);
    if (!isValid(id)) throw bad_id(id);
    return _items[id];
}

void DataManager::remove(id_t id) {
    std::lock_guard<std::mutex> g(_m);
    if (!isValid(id)) throw bad_id(id);
    _items[id] = nullptr;
}

void DataManager::remove(std::vector<id_t> const& ids) {
    for (id_t id : ids) remove(id);
}

void DataManager::removeAll() {
    std::lock_guard<std::mutex> g(_m);
    _items.clear();
}

}  // namespace as

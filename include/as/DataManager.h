/*
 * Host-side registry of the data a request's Common refers to by id, with the
 * reference's interface and error strings (server/DataManager.h:22-75,
 * server/DataManager.cpp:19-60).  The items hold plain arrays in the layouts
 * of the C-ABI descriptors (include/attract_hip.h); a caller that loads its
 * proteins, grids and tables with the reference's readers converts them with
 * the accessors of those objects (the adapter is a few lines per type).
 */
#ifndef AS_DATAMANAGER_H_
#define AS_DATAMANAGER_H_

#include <memory>
#include <mutex>
#include <vector>

#include "Types_6D.h"
#include "../attract_hip.h"

namespace as {

class DataItem {
public:
    virtual ~DataItem() {}
};

/* Protein (model/Protein.h:38-167): pivotized coordinates, raw types (the
 * neighbour-list term), mapped types (grid term), charges, modes in the
 * [x: n*M][y][z] layout with element i*M+m, mode force constants. */
class ProteinData : public DataItem {
public:
    unsigned n = 0, numModes = 0;
    std::vector<float> xyz;          // 3n, SoA
    std::vector<unsigned> type, mappedType;
    std::vector<float> charge;
    std::vector<float> modes;        // 3*n*numModes
    std::vector<float> modeForce;    // numModes
    att_protein_desc desc() const;
};

/* GridUnion (model/GridUnion.h): inner and outer interpolation grids and,
 * for a receptor, the neighbour-list grid. */
class GridUnionData : public DataItem {
public:
    struct Grid {
        unsigned nx = 0, ny = 0, nz = 0, numGrids = 0;
        float ox = 0, oy = 0, oz = 0, spacing = 0;
        std::vector<float> data;     // numGrids*nz*ny*nx float4 {fx,fy,fz,E}
    } inner, outer;
    bool hasNL = false;
    struct NL {
        unsigned nx = 0, ny = 0, nz = 0;
        float ox = 0, oy = 0, oz = 0, spacing = 0, plateau = 0;
        std::vector<uint32_t> desc;  // nz*ny*nx {numEl, idx}
        std::vector<uint32_t> list;
    } nl;
    att_grid_desc innerDesc() const;
    att_grid_desc outerDesc() const;
    att_nlgrid_desc nlDesc() const;
};

/* ParamTable (model/ParamTable.h:18-101) after the derivation of
 * readFile.cpp:479-513: T*T {rc, ac, ipon, rmin2, emin}, index
 * [recType-1][ligType-1]. */
class ParamTableData : public DataItem {
public:
    unsigned numTypes = 0;
    int potShape = 8;
    std::vector<att_pair_params> table;
};

/* SimParam (model/SimParam.h:39-54). */
class SimParamData : public DataItem {
public:
    int dielec = 1;          // 0 constant, 1 variable
    float epsilon = 15.f;
    float ffelec = 0.f;      // 332.053986 / epsilon
};

class DataManager {
public:
    id_t add(std::shared_ptr<DataItem> dataItem);
    std::vector<id_t> add(std::vector<std::shared_ptr<DataItem>> const& items);
    std::shared_ptr<DataItem> get(id_t id);
    void remove(id_t id);
    void remove(std::vector<id_t> const& ids);
    void removeAll();

private:
    bool isValid(id_t id) const;
    mutable std::mutex _m;
    std::vector<std::shared_ptr<DataItem>> _items;
};

}  // namespace as

#endif

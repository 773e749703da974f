/*
 * Request/result types of the scoring service, layout-identical to the
 * reference's (model/Types_6D.h:31-54, model/Types_6D_Modes.h:33-60,
 * model/GenericTypes.h, server/publicTypes.h), so that buffers written for the
 * reference pass through unchanged.  The structs are also layout-identical to
 * the C-ABI's att_dof6 / att_result6 / att_dof6m / att_result6m.
 */
#ifndef AS_TYPES_6D_H_
#define AS_TYPES_6D_H_

#include <cstddef>

/* The reference declares GenericTypes at global scope. */
template <typename InputType, typename CommonType, typename ResultType>
struct GenericTypes {
    using input_t = InputType;
    using common_t = CommonType;
    using result_t = ResultType;
};

namespace as {

using id_t = unsigned;
using deviceId_t = unsigned;
using workerId_t = unsigned;

constexpr unsigned MODES_MAX_NUMBER = 10;

template <typename REAL>
struct Vec3 {
    REAL x, y, z;
    Vec3() = default;
    Vec3(REAL a, REAL b, REAL c) : x(a), y(b), z(c) {}
};

template <typename REAL>
struct DOF_6D {
    Vec3<REAL> pos;
    Vec3<REAL> ang;
};

template <typename REAL>
struct Result_6D {
    REAL E;
    Vec3<REAL> pos;  // sum of bead forces = -dE/dpos
    Vec3<REAL> ang;  // torque contracted with dR/dangle = -dE/dang
};

struct Common {
    id_t gridId;
    id_t ligId;
    id_t recId;
    id_t tableId;
    id_t paramsId;
};

template <typename REAL>
using Types_6D = GenericTypes<DOF_6D<REAL>, Common, Result_6D<REAL>>;

template <typename REAL>
struct DOF_6D_Modes {
    DOF_6D<REAL> _6D;
    REAL modesRec[MODES_MAX_NUMBER];
    REAL modesLig[MODES_MAX_NUMBER];
};

template <typename REAL>
struct Result_6D_Modes {
    Result_6D<REAL> _6D;
    REAL modesRec[MODES_MAX_NUMBER];
    REAL modesLig[MODES_MAX_NUMBER];
};

struct Common_Modes {
    id_t gridIdRec;
    id_t gridIdLig;
    id_t ligId;
    id_t recId;
    id_t tableId;
    id_t paramsId;
};

template <typename REAL>
using Types_6D_Modes = GenericTypes<DOF_6D_Modes<REAL>, Common_Modes, Result_6D_Modes<REAL>>;

static_assert(sizeof(DOF_6D<float>) == 24 && sizeof(Result_6D<float>) == 28, "6D layout");
static_assert(sizeof(DOF_6D_Modes<float>) == 104 && sizeof(Result_6D_Modes<float>) == 108,
              "6D modes layout");

}  // namespace as

#endif

// The reference's input and output files on the host (include/attract_formats.h).
//
// Numbers are extracted exactly as the reference's readers do (iostream
// extraction into float, double or unsigned; stof/stoi), and every conversion
// keeps the reference's float arithmetic in its order, so the arrays equal the
// reference's bit for bit.
#include "attract_formats.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string t_err;

att_status err(att_status st, const std::string& msg) {
    t_err = msg;
    return st;
}

std::vector<std::string> split_ws(const std::string& line) {
    std::istringstream iss(line);
    std::vector<std::string> out;
    std::string t;
    while (iss >> t) out.push_back(t);
    return out;
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

struct Rot {
    float m[9];
};

// euler2rotmat (model/matrixFunctions.h:38-77) in float.
Rot euler2rot(float phi, float ssi, float rot) {
    const float cSSI = std::cos(ssi), cPHI = std::cos(phi), sSSI = std::sin(ssi), sPHI = std::sin(phi);
    const float cSSI_cPHI = cSSI * cPHI, cSSI_sPHI = cSSI * sPHI;
    const float sSSI_cPHI = sSSI * cPHI, sSSI_sPHI = sSSI * sPHI;
    const float cROT = std::cos(rot), sROT = std::sin(rot);
    Rot r;
    r.m[0] = cROT * cSSI_cPHI + sROT * sPHI;
    r.m[1] = sROT * cSSI_cPHI - cROT * sPHI;
    r.m[2] = sSSI_cPHI;
    r.m[3] = cROT * cSSI_sPHI - sROT * cPHI;
    r.m[4] = sROT * cSSI_sPHI + cROT * cPHI;
    r.m[5] = sSSI_sPHI;
    r.m[6] = -cROT * sSSI;
    r.m[7] = -sROT * sSSI;
    r.m[8] = cSSI;
    return r;
}

// RotMat::operator* (commons/RotMat.h:87-97): sums start from zero.
Rot mul(const Rot& a, const Rot& b) {
    Rot o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.f;
            for (int k = 0; k < 3; ++k) s += a.m[i * 3 + k] * b.m[k * 3 + j];
            o.m[i * 3 + j] = s;
        }
    return o;
}

// rotmat2euler (model/matrixFunctions.h:131-166), including its test of
// abs(m[8] >= 0.9999), which is true only for m[8] >= 0.9999.
void rot2euler(const Rot& r, float& phi, float& ssi, float& rot) {
    const float* m = r.m;
    phi = std::atan2(m[5], m[2]);
    ssi = (float)std::acos(std::fmin(std::fmax((double)m[8], -1.0), 1.0));
    rot = std::atan2(-m[7], -m[6]);
    if ((double)m[8] >= 0.9999) {
        phi = 0.f;
        if ((double)std::fabs(m[0]) >= 0.9999) {
            rot = m[0] < 0.f ? (float)M_PI : 0.f;
            ssi = m[8] < 0.f ? (float)M_PI : 0.f;
        } else {
            if (m[8] < 0.f) {
                ssi = (float)M_PI;
                rot = -std::acos(-m[0]);
            } else {
                ssi = 0.f;
                rot = std::acos(m[0]);
            }
        }
        if (m[1] < 0) rot = (float)(rot * -1.0);
    }
}

}  // namespace

struct att_body {
    unsigned n = 0, M = 0;
    std::vector<float> xyz;  // SoA
    std::vector<unsigned> type, mapped;
    std::vector<float> charge, modes, modeForce;
    float pivot[3] = {0.f, 0.f, 0.f};
};

struct DofRec {
    double ang[3], pos[3];
    double extra[10];
    unsigned nExtra;
};

struct att_dof_file {
    att_dof_info info{};
    std::vector<std::vector<DofRec>> mol;  // [molecule][pose]
};

extern "C" {

const char* att_formats_last_error(void) { return t_err.c_str(); }

att_status att_body_read_pdb(const char* path, att_body** out) {
    if (!path || !out) return err(ATT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::ifstream file(path);
    if (!file.is_open()) return err(ATT_ERR_INVALID, std::string("Failed to open file ") + path);
    std::vector<float> X, Y, Z, q;
    std::vector<unsigned> t;
    static const int idx[5][2] = {{30, 38}, {38, 46}, {46, 54}, {54, 59}, {59, 67}};
    std::string line;
    size_t lineNo = 1;
    std::getline(file, line);
    while (!file.eof()) {
        if (!line.empty()) {
            std::string tok[5];
            for (int i = 0; i < 5; ++i) {
                if ((size_t)idx[i][0] > line.size())
                    return err(ATT_ERR_INVALID, std::string(path) + ": line " + std::to_string(lineNo) +
                                                    " ends before column " + std::to_string(idx[i][0] + 1));
                tok[i] = line.substr(idx[i][0], idx[i][1] - idx[i][0]);
                tok[i].erase(std::remove_if(tok[i].begin(), tok[i].end(), ::isspace), tok[i].end());
                if (tok[i].empty())
                    return err(ATT_ERR_INVALID, std::string(path) + ": line " + std::to_string(lineNo) +
                                                    " has an empty column " + std::to_string(i + 1));
            }
            float x = 0, y = 0, z = 0, c = 0;
            unsigned ty = 0;
            std::stringstream(tok[0]) >> x;
            std::stringstream(tok[1]) >> y;
            std::stringstream(tok[2]) >> z;
            std::stringstream(tok[3]) >> ty;
            std::stringstream(tok[4]) >> c;
            X.push_back(x);
            Y.push_back(y);
            Z.push_back(z);
            t.push_back(ty);
            q.push_back(c);
        }
        std::getline(file, line);
        ++lineNo;
    }
    std::unique_ptr<att_body> b(new att_body);
    b->n = (unsigned)t.size();
    b->xyz = X;
    b->xyz.insert(b->xyz.end(), Y.begin(), Y.end());
    b->xyz.insert(b->xyz.end(), Z.begin(), Z.end());
    b->type = t;
    b->mapped = t;
    b->charge = q;
    *out = b.release();
    return ATT_OK;
}

att_status att_body_read_modes(att_body* b, const char* path, unsigned numModes) {
    if (!b || !path) return err(ATT_ERR_INVALID, "null argument");
    if (numModes == 0 || numModes > ATT_MODES_MAX) return err(ATT_ERR_INVALID, "number of modes must be 1..10");
    std::ifstream file(path);
    if (!file.is_open()) return err(ATT_ERR_INVALID, std::string("error while opening file ") + path);
    const int M = (int)numModes, n = (int)b->n;
    std::vector<float> modes((size_t)3 * M * n, 0.f), modeVal(M, 0.f), eigVal(M, 0.f);
    std::vector<bool> seen(M, false);
    bool changeMode = true, isData = false;
    int idxPos = 0, modeIdx = 0;
    std::string line;
    try {
        while (std::getline(file, line)) {
            if (line.empty()) continue;
            const std::vector<std::string> tok = split_ws(line);
            if (idxPos < n && !changeMode) isData = true;
            else if (idxPos == n) {
                isData = false;
                idxPos = 0;
            }
            if (changeMode && tok.size() > 1) {
                ++modeIdx;
                if (modeIdx == std::stoi(tok.at(0)) && modeIdx <= M) {
                    const float e = std::stof(tok.at(1));
                    modeVal[modeIdx - 1] = e * std::stof(tok.at(1));
                    eigVal[modeIdx - 1] = 0.f;
                    seen[modeIdx - 1] = true;
                    changeMode = false;
                }
            }
            if (isData && tok.size() == 3) {
                if (modeIdx < 1 || modeIdx > M)
                    return err(ATT_ERR_INVALID, std::string(path) + ": data line outside a wanted mode");
                const float x = std::stof(tok.at(0)), y = std::stof(tok.at(1)), z = std::stof(tok.at(2));
                eigVal[modeIdx - 1] += x * x + y * y + z * z;
                modes[(size_t)M * idxPos + modeIdx - 1] = std::stof(tok.at(0));
                modes[(size_t)n * M + (size_t)M * idxPos + modeIdx - 1] = std::stof(tok.at(1));
                modes[(size_t)2 * n * M + (size_t)M * idxPos + modeIdx - 1] = std::stof(tok.at(2));
                ++idxPos;
                if (idxPos == n) changeMode = true;
            }
        }
    } catch (const std::exception& e) {
        return err(ATT_ERR_INVALID, std::string(path) + ": unreadable number (" + e.what() + ")");
    }
    for (int m = 0; m < M; ++m)
        if (!seen[m]) return err(ATT_ERR_INVALID, std::string(path) + ": mode " + std::to_string(m + 1) + " missing");
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < M; ++m) {
            const float s = std::sqrt(eigVal[m]);
            modes[(size_t)M * i + m] /= s;
            modes[(size_t)n * M + (size_t)M * i + m] /= s;
            modes[(size_t)2 * n * M + (size_t)M * i + m] /= s;
        }
    b->M = numModes;
    b->modes = modes;
    b->modeForce = modeVal;
    return ATT_OK;
}

att_status att_body_pivotize(att_body* b, const float* pivot, float* pivot_out) {
    if (!b) return err(ATT_ERR_INVALID, "null body");
    float p[3];
    if (pivot) {
        p[0] = pivot[0]; p[1] = pivot[1]; p[2] = pivot[2];
    } else {
        // auto_pivotize: centroid of the original coordinates, float sums
        float s[3] = {0.f, 0.f, 0.f};
        for (unsigned i = 0; i < b->n; ++i)
            for (int k = 0; k < 3; ++k) s[k] += b->xyz[(size_t)k * b->n + i] + b->pivot[k];
        for (int k = 0; k < 3; ++k) p[k] = s[k] / (float)(double)b->n;
    }
    // undo the previous pivot, then apply the new one (Protein.tpp:130-167)
    for (int k = 0; k < 3; ++k) {
        float* c = b->xyz.data() + (size_t)k * b->n;
        if (b->pivot[0] != 0.f || b->pivot[1] != 0.f || b->pivot[2] != 0.f)
            for (unsigned i = 0; i < b->n; ++i) c[i] += b->pivot[k];
    }
    b->pivot[0] = p[0]; b->pivot[1] = p[1]; b->pivot[2] = p[2];
    if (p[0] != 0.f || p[1] != 0.f || p[2] != 0.f)
        for (int k = 0; k < 3; ++k) {
            float* c = b->xyz.data() + (size_t)k * b->n;
            for (unsigned i = 0; i < b->n; ++i) c[i] -= p[k];
        }
    if (pivot_out) {
        pivot_out[0] = p[0]; pivot_out[1] = p[1]; pivot_out[2] = p[2];
    }
    return ATT_OK;
}

att_status att_body_map_types(att_body* b, const unsigned* alphabet, unsigned n_alphabet) {
    if (!b || (n_alphabet && !alphabet)) return err(ATT_ERR_INVALID, "null argument");
    for (unsigned i = 0; i < b->n; ++i) {
        const unsigned t = b->type[i];
        if (n_alphabet == 0) {
            b->mapped[i] = t;
            continue;
        }
        if (t == 0) {
            b->mapped[i] = 0;
            continue;
        }
        // the last occurrence wins, as repeated std::map assignments do
        int found = -1;
        for (unsigned k = 0; k < n_alphabet; ++k)
            if (alphabet[k] == t) found = (int)k;
        if (found < 0) return err(ATT_ERR_INVALID, "bead type " + std::to_string(t) + " is not in the grid alphabet");
        b->mapped[i] = t == 99 ? 0u : (unsigned)found + 1;
    }
    return ATT_OK;
}

att_status att_body_desc(const att_body* b, att_protein_desc* out) {
    if (!b || !out) return err(ATT_ERR_INVALID, "null argument");
    *out = att_protein_desc{b->n, b->M, b->xyz.data(), b->type.data(), b->mapped.data(), b->charge.data(),
                            b->M ? b->modes.data() : nullptr, b->M ? b->modeForce.data() : nullptr};
    return ATT_OK;
}

void att_body_free(att_body* b) { delete b; }

att_status att_read_alphabet(const char* path, unsigned* out, unsigned capacity, unsigned* n) {
    if (!path || !n) return err(ATT_ERR_INVALID, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) return err(ATT_ERR_INVALID, std::string("Error: Failed to open file ") + path);
    std::vector<unsigned> types;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int key = 0;
        if (!(ls >> key) || key < 0)
            return err(ATT_ERR_INVALID, std::string(path) + ": line " +
                                            std::to_string(types.size() + 1) +
                                            " is not a non-negative bead type");
        types.push_back((unsigned)key);
    }
    *n = (unsigned)types.size();
    if (!out) return ATT_OK;
    if (types.size() > capacity) return err(ATT_ERR_INVALID, "alphabet capacity too small");
    std::copy(types.begin(), types.end(), out);
    return ATT_OK;
}

att_status att_read_params(const char* path, att_pair_params* table, unsigned capacity,
                           unsigned* numTypes, int* potShape, float* swiOn, float* swiOff) {
    if (!path || !numTypes) return err(ATT_ERR_INVALID, "null argument");
    std::ifstream in(path, std::ios::in);
    if (in.fail()) return err(ATT_ERR_INVALID, std::string("Error: Failed to open file: ") + path);
    unsigned shape = 0, T = 0;
    float on = 0.f, off = 0.f;
    in >> shape >> T >> on >> off;
    if (!in) return err(ATT_ERR_INVALID, std::string(path) + ": unreadable header");
    *numTypes = T;
    if (potShape) *potShape = (int)shape;
    if (swiOn) *swiOn = on;
    if (swiOff) *swiOff = off;
    if (shape != 8 && shape != 12)
        return err(ATT_ERR_INVALID, "Unknown potential shape " + std::to_string(shape) + " in file: " + path);
    if (!table) return ATT_OK;  // size query
    if ((size_t)T * T > capacity) return err(ATT_ERR_INVALID, "table capacity below numTypes^2");
    const size_t TT = (size_t)T * T;
    std::vector<float> rbc(TT), abc(TT);
    std::vector<int> ipon(TT);
    for (size_t k = 0; k < TT; ++k) in >> rbc[k];
    for (size_t k = 0; k < TT; ++k) in >> abc[k];
    for (size_t k = 0; k < TT; ++k) in >> ipon[k];
    if (!in) return err(ATT_ERR_INVALID, std::string(path) + ": fewer than 3 x numTypes^2 values");
    for (size_t k = 0; k < TT; ++k) {
        att_pair_params& p = table[k];
        const float r = rbc[k], A = abc[k];
        p.ipon = ipon[k];
        if (shape == 8) {
            p.rc = (float)(A * std::pow((double)r, 8.0));
            p.ac = (float)(A * std::pow((double)r, 6.0));
            if (p.ac > 0 && p.rc > 0) {
                p.emin = (float)(-27.0f * std::pow((double)p.ac, 4.0) / (256.0f * std::pow((double)p.rc, 3.0)));
                p.rmin2 = 4.0f * p.rc / (3.0f * p.ac);
            } else {
                p.emin = 0.f;
                p.rmin2 = 0.f;
            }
        } else {
            p.rc = (float)(A * std::pow((double)r, 12.0));
            p.ac = (float)(A * std::pow((double)r, 6.0));
            p.emin = (float)(-0.25 * A);
            p.rmin2 = 1.25992105f * r * r;
        }
    }
    return ATT_OK;
}

att_status att_dof_read(const char* path, att_dof_file** out) {
    if (!path || !out) return err(ATT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<att_dof_file> f(new att_dof_file);
    att_dof_info& I = f->info;
    const std::string fmt = std::string("Error reading file ") + path + ". DOF format is not supported. ";
    {
        // header: pivots and centering (readDOFHeader)
        std::ifstream file(path);
        if (!file.is_open()) return err(ATT_ERR_INVALID, std::string("Error: Failed to open file ") + path);
        bool pivDone = false, recDone = false, ligDone = false;
        std::string line;
        while (!file.eof()) {
            if (pivDone && recDone && ligDone) break;
            std::getline(file, line);
            if (starts_with(line, "##")) continue;
            const std::vector<std::string> tok = split_ws(line);
            if (tok.empty()) continue;
            if (tok[0] == "#pivot") {
                if (pivDone) return err(ATT_ERR_INVALID, fmt + "Incorrect pivot definition.");
                if (tok.size() > 1 && tok[1] == "auto") {
                    I.autoPivot = 1;
                    pivDone = true;
                    continue;
                }
                if (tok.size() != 5) return err(ATT_ERR_INVALID, fmt + "Incorrect pivot definition.");
                float p[3];
                for (int k = 0; k < 3; ++k) std::stringstream(tok[2 + k]) >> p[k];
                if (I.numPivots < 2)
                    for (int k = 0; k < 3; ++k) I.pivots[I.numPivots][k] = p[k];
                ++I.numPivots;
                continue;
            }
            if (tok[0] == "#centered" && (!recDone || !ligDone)) {
                pivDone = true;
                if (tok.size() < 3) return err(ATT_ERR_INVALID, fmt + "Centering definition invalid.");
                int v;
                if (tok[2] == "true") v = 1;
                else if (tok[2] == "false") v = 0;
                else
                    return err(ATT_ERR_INVALID, fmt + (tok[1] == "receptor:" ? "Receptor" : "Ligand") +
                                                    " centering definition invalid.");
                if (tok[1] == "receptor:") {
                    I.centeredReceptor = v;
                    recDone = true;
                } else if (tok[1] == "ligand:" || tok[1] == "ligands:") {
                    I.centeredLigands = v;
                    ligDone = true;
                } else {
                    return err(ATT_ERR_INVALID, fmt + "Centering definition invalid.");
                }
            }
        }
        if (!recDone || !ligDone) return err(ATT_ERR_INVALID, fmt + "Centering definitions missing.");
    }
    // pose blocks (readDOF): lines up to the next '#' line, one per molecule
    std::ifstream file(path);
    std::string line;
    size_t lineNo = 0;
    unsigned block = 0;
    while (!file.eof()) {
        std::getline(file, line);
        ++lineNo;
        if (starts_with(line, "#")) continue;
        unsigned i = 0;
        while (!starts_with(line, "#") && !file.eof()) {
            if (block == 0) f->mol.emplace_back();
            if (i >= f->mol.size())
                return err(ATT_ERR_INVALID, fmt + "More molecules at line " + std::to_string(lineNo) +
                                                " than in the first pose.");
            DofRec d{};
            std::stringstream s(line);
            s >> d.ang[0] >> d.ang[1] >> d.ang[2] >> d.pos[0] >> d.pos[1] >> d.pos[2];
            if (!s) return err(ATT_ERR_INVALID, fmt + "Fewer than 6 values at line " + std::to_string(lineNo) + ".");
            double v;
            while (s >> v) {
                if (d.nExtra == 10)
                    return err(ATT_ERR_INVALID, fmt + "More than 10 mode values at line " + std::to_string(lineNo) + ".");
                d.extra[d.nExtra++] = v;
            }
            f->mol[i].push_back(d);
            ++i;
            std::getline(file, line);
            ++lineNo;
        }
        if (i != f->mol.size())
            return err(ATT_ERR_INVALID, fmt + "The DOF definition is incomplete at #" + std::to_string(block) +
                                            " (line " + std::to_string(lineNo) + ")");
        ++block;
    }
    I.numMolecules = (unsigned)f->mol.size();
    I.numPoses = f->mol.empty() ? 0 : f->mol[0].size();
    *out = f.release();
    return ATT_OK;
}

att_status att_dof_get_info(const att_dof_file* f, att_dof_info* info) {
    if (!f || !info) return err(ATT_ERR_INVALID, "null argument");
    *info = f->info;
    return ATT_OK;
}

att_status att_dof_values(const att_dof_file* f, unsigned molecule, size_t pose, double* values,
                          unsigned* count) {
    if (!f || !values || !count) return err(ATT_ERR_INVALID, "null argument");
    if (molecule >= f->mol.size() || pose >= f->mol[molecule].size())
        return err(ATT_ERR_INVALID, "no such molecule or pose in the DOF file");
    const DofRec& d = f->mol[molecule][pose];
    for (int k = 0; k < 3; ++k) {
        values[k] = d.ang[k];
        values[3 + k] = d.pos[k];
    }
    for (unsigned k = 0; k < d.nExtra; ++k) values[6 + k] = d.extra[k];
    *count = 6 + d.nExtra;
    return ATT_OK;
}

att_status att_dof_ligand(const att_dof_file* f, const float* pivotRec, const float* pivotLig,
                          int modes, unsigned numModesRec, unsigned numModesLig, void* out) {
    if (!f || !out) return err(ATT_ERR_INVALID, "null argument");
    if (f->info.numMolecules != 2)
        return err(ATT_ERR_INVALID, "DOF-file contains definitions for more than two molecules. "
                                    "Multi-body docking is not supported.");
    if (numModesRec > ATT_MODES_MAX || numModesLig > ATT_MODES_MAX) return err(ATT_ERR_INVALID, "more than 10 modes");
    float pr[3], pl[3];
    if (f->info.autoPivot) {
        if (f->info.numPivots)
            return err(ATT_ERR_INVALID, "Auto pivot specified, but explicitly defined pivots available.");
        if (!pivotRec || !pivotLig) return err(ATT_ERR_INVALID, "auto pivot: the bodies' pivots are needed");
        for (int k = 0; k < 3; ++k) { pr[k] = pivotRec[k]; pl[k] = pivotLig[k]; }
    } else {
        if (f->info.numPivots != 2)
            return err(ATT_ERR_INVALID, "No auto pivot specified, but number of defined pivots is incorrect.");
        for (int k = 0; k < 3; ++k) { pr[k] = f->info.pivots[0][k]; pl[k] = f->info.pivots[1][k]; }
    }
    const size_t n = f->mol[1].size();
    if (f->mol[0].size() != n) return err(ATT_ERR_INVALID, "receptor and ligand pose counts differ");
    float* o = static_cast<float*>(out);
    const size_t stride = modes ? 26 : 6;
    for (size_t p = 0; p < n; ++p) {
        const DofRec& R = f->mol[0][p];
        const DofRec& L = f->mol[1][p];
        // DOFConverter: double -> float
        float rpos[3], rang[3], lpos[3], lang[3];
        for (int k = 0; k < 3; ++k) {
            rpos[k] = (float)R.pos[k]; rang[k] = (float)R.ang[k];
            lpos[k] = (float)L.pos[k]; lang[k] = (float)L.ang[k];
        }
        // transformDOF_glob2rec (model/DOFTransform.cpp:36-84)
        if (!f->info.centeredLigands)
            for (int k = 0; k < 3; ++k) lpos[k] += pl[k];
        if (!f->info.centeredReceptor && (pr[0] != 0.f || pr[1] != 0.f || pr[2] != 0.f))
            for (int k = 0; k < 3; ++k) lpos[k] -= pr[k];
        if (rpos[0] != 0.f || rpos[1] != 0.f || rpos[2] != 0.f)
            for (int k = 0; k < 3; ++k) lpos[k] = lpos[k] - rpos[k];
        if (rang[0] != 0.f || rang[1] != 0.f || rang[2] != 0.f) {
            const Rot inv = euler2rot(-rang[0], -rang[1], -rang[2]);
            const Rot ml = euler2rot(lang[0], lang[1], lang[2]);
            float v[3];
            for (int r = 0; r < 3; ++r)
                v[r] = inv.m[3 * r] * lpos[0] + inv.m[3 * r + 1] * lpos[1] + inv.m[3 * r + 2] * lpos[2];
            for (int k = 0; k < 3; ++k) lpos[k] = v[k];
            rot2euler(mul(inv, ml), lang[0], lang[1], lang[2]);
        }
        float* d = o + stride * p;
        for (int k = 0; k < 3; ++k) {
            d[k] = lpos[k];
            d[3 + k] = lang[k];
        }
        if (modes) {
            // DOFConverter_Modes: missing or NaN amplitudes are 0
            for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
                d[6 + m] = (m < numModesRec && m < R.nExtra && !std::isnan(R.extra[m])) ? (float)R.extra[m] : 0.f;
                d[16 + m] = (m < numModesLig && m < L.nExtra && !std::isnan(L.extra[m])) ? (float)L.extra[m] : 0.f;
            }
        }
    }
    return ATT_OK;
}

void att_dof_free(att_dof_file* f) { delete f; }

att_status att_format_results(const void* results, size_t n, int modes, unsigned numModesRec,
                              unsigned numModesLig, char* buf, size_t cap, size_t* len) {
    if ((!results && n) || !len) return err(ATT_ERR_INVALID, "null argument");
    if (numModesRec > ATT_MODES_MAX || numModesLig > ATT_MODES_MAX) return err(ATT_ERR_INVALID, "more than 10 modes");
    std::ostringstream s;
    const float* r = static_cast<const float*>(results);
    const size_t stride = modes ? 27 : 7;
    for (size_t p = 0; p < n; ++p) {
        const float* e = r + stride * p;
        const float E = e[0];
        s.setf(std::ios::scientific);
        s.precision(8);
        s << " Energy: " << E << std::endl;
        s.unsetf(std::ios::scientific);
        s.setf(std::ios::fixed);
        s.precision(3);
        s << std::setw(12) << E << std::endl;
        s.unsetf(std::ios::fixed);
        s.setf(std::ios::scientific);
        s.precision(8);
        const int w = 20;
        if (!modes) {
            s << " Gradients: " << std::setw(w) << e[4] << std::setw(w) << e[5] << std::setw(w) << e[6]
              << std::setw(w) << e[1] << std::setw(w) << e[2] << std::setw(w) << e[3];
        } else {
            for (int k = 0; k < 6; ++k) s << std::setw(w) << 0;
            for (unsigned m = 0; m < numModesRec; ++m) s << std::setw(w) << e[7 + m];
            s << std::setw(w) << e[4] << std::setw(w) << e[5] << std::setw(w) << e[6] << std::setw(w) << e[1]
              << std::setw(w) << e[2] << std::setw(w) << e[3];
            for (unsigned m = 0; m < numModesLig; ++m) s << std::setw(w) << e[17 + m];
        }
        s.unsetf(std::ios::scientific);
        s.precision(6);
        s << std::endl;
    }
    const std::string str = s.str();
    *len = str.size();
    if (buf) std::memcpy(buf, str.data(), std::min(cap, str.size()));
    return ATT_OK;
}

}  // extern "C"

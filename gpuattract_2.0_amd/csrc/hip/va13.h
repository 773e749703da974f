// VA13 variable-metric minimizer (apps/em/minfor.f:1-182, factorized Hessian
// updates of apps/em/mc11a.f:1-143) as a resumable state machine: a pose's
// minimizer runs until it needs the objective at a new point, writes that
// point and returns; after the point has been scored it resumes with the value
// and gradient.  This is how many poses minimize at once on the device, one
// scoring launch per round of requests.  Double precision throughout, the
// operations in the order of the Fortran source.
//
// The Hessian approximation is L D L^T in packed storage: for i = 0..n-1, D_i
// followed by L_{i+1..n-1, i}.  A pose's vectors and factors live in one
// array, element e of pose p at st[e * stride + p] (coalesced over poses).
#pragma once

#include <hip/hip_runtime.h>

namespace att {

constexpr int kVaMaxDim = 26;  // 6 rigid DOFs + 10 + 10 modes

enum VaVec : int { kVaX = 0, kVaXaa, kVaXbb, kVaG, kVaGa, kVaGb, kVaD, kVaW, kVaNumVec };

// Scalar state of one pose's minimizer.
struct VaState {
    double gesa, fa, fb, dff, dga, dgb, c, stmin, stepbd, steplb, fmn, gmin, step;
    int isfv, itr, nfun, ir, pc, pad;
};

enum VaPc : int { kVaFresh = 0, kVaAtStart = 1, kVaInLine = 2, kVaFinished = 3 };

// One pose's vectors: vector v, element k at st[(v * nmax + k) * stride];
// the packed factors after the vectors.
struct VaView {
    double* st;
    size_t stride;
    int nmax;
    __host__ __device__ double& v(int vec, int k) const {
        return st[((size_t)vec * nmax + k) * stride];
    }
    __host__ __device__ double& h(int k) const {
        return st[((size_t)kVaNumVec * nmax + k) * stride];
    }
};

__host__ __device__ inline size_t va_elements(int nmax) {
    return (size_t)kVaNumVec * nmax + (size_t)nmax * (nmax + 1) / 2;
}

__host__ __device__ inline double va_min(double a, double b) { return b < a ? b : a; }
__host__ __device__ inline double va_max(double a, double b) { return b > a ? b : a; }

// H <- H + sig z z^T on the factors (mc11a): with mk = 1 the vector W holds
// L^-1 z from the last solve.  `z` and `w` name vectors of the view and may
// be the same vector.
__host__ __device__ inline void va_update(const VaView& V, int n, int z, double sig, int w,
                                          int& ir, int mk, double eps) {
    if (n <= 1) {
        V.h(0) = V.h(0) + sig * V.v(z, 0) * V.v(z, 0);
        ir = 1;
        if (V.h(0) > 0.0) return;
        V.h(0) = 0.0;
        ir = 0;
        return;
    }
    double ti = 0.0, tim = 0.0;
    int mm = 0;
    int pos = 0;
    bool plus = sig > 0.0;
    if (!plus) {
        if (sig == 0.0 || ir == 0) return;
        ti = 1.0 / sig;
        if (mk != 0) {
            for (int i = 0; i < n; ++i) {
                const double a = V.h(pos);
                if (a != 0.0) ti = ti + V.v(w, i) * V.v(w, i) / a;
                pos += n - i;
            }
        } else {
            for (int i = 0; i < n; ++i) V.v(w, i) = V.v(z, i);
            for (int i = 0; i < n; ++i) {
                const double val = V.v(w, i);
                const double a = V.h(pos);
                if (!(a > 0.0)) {
                    V.v(w, i) = 0.0;
                    pos += n - i;
                    continue;
                }
                ti = ti + val * val / a;
                for (int j = i + 1; j < n; ++j) V.v(w, j) = V.v(w, j) - val * V.h(++pos);
                ++pos;
            }
        }
        bool recur = true;
        if (ir <= 0) {
            ti = 0.0;
            ir = -ir - 1;
        } else if (ti > 0.0) {
            ti = eps / sig;
            if (eps == 0.0) ir = ir - 1;
        } else if (mk <= 1) {
            plus = true;
            recur = false;
        }
        if (recur) {
            // partial sums from the last variable backwards
            mm = 1;
            tim = ti;
            for (int i = 1; i <= n; ++i) {
                const int j = n - i;
                pos -= i;
                const double a = V.h(pos);
                if (a != 0.0) tim = ti - V.v(w, j) * V.v(w, j) / a;
                V.v(w, j) = ti;
                ti = tim;
            }
        }
    }
    if (plus) {
        mm = 0;
        tim = 1.0 / sig;
    }
    pos = 0;
    for (int i = 0; i < n; ++i) {
        const double val = V.v(z, i);
        const double a = V.h(pos);
        if (!(a > 0.0)) {
            if (ir > 0 || sig < 0.0 || val == 0.0) {
                ti = tim;
                pos += n - i;
                continue;
            }
            ir = 1 - ir;
            V.h(pos) = val * val / tim;
            for (int j = i + 1; j < n; ++j) V.h(++pos) = V.v(z, j) / val;
            return;
        }
        const double al = val / a;
        ti = mm <= 0 ? tim + val * al : V.v(w, i);
        const double r = ti / tim;
        V.h(pos) = a * r;
        if (r == 0.0 || i == n - 1) break;
        const double b = al / ti;
        if (r > 4.0) {
            const double gm = tim / ti;
            for (int j = i + 1; j < n; ++j) {
                const double y = V.h(++pos);
                V.h(pos) = b * V.v(z, j) + y * gm;
                V.v(z, j) = V.v(z, j) - val * y;
            }
        } else {
            for (int j = i + 1; j < n; ++j) {
                ++pos;
                V.v(z, j) = V.v(z, j) - val * V.h(pos);
                V.h(pos) = V.h(pos) + b * V.v(z, j);
            }
        }
        tim = ti;
        ++pos;
    }
    if (ir < 0) ir = -ir;
}

// z <- H^-1 z (mc11e); W receives L^-1 z.
__host__ __device__ inline void va_solve(const VaView& V, int n, int z) {
    V.v(kVaW, 0) = V.v(z, 0);
    if (n <= 1) {
        V.v(z, 0) = V.v(z, 0) / V.h(0);
        return;
    }
    int pos = 0;
    for (int i = 1; i < n; ++i) {
        pos = i;
        double val = V.v(z, i);
        for (int j = 0; j < i; ++j) {
            val = val - V.h(pos) * V.v(z, j);
            pos += n - 1 - j;
        }
        V.v(kVaW, i) = val;
        V.v(z, i) = val;
    }
    V.v(z, n - 1) = V.v(z, n - 1) / V.h(pos);
    for (int back = 2; back <= n; ++back) {
        const int i = n - back;
        int q = pos - back;
        double val = V.v(z, i) / V.h(q);
        pos = q;
        for (int j = i + 1; j < n; ++j) val = val - V.h(++q) * V.v(z, j);
        V.v(z, i) = val;
    }
}

// Advances one pose's minimizer.  On kVaFresh the start point must be in X.
// f, g: the objective and its gradient at the point last requested (ignored
// when fresh).  Returns true with the next point in `req`, or false when the
// minimizer has finished (the last requested point is then its result).
__host__ __device__ inline bool va_resume(VaState& s, const VaView& V, int n, int ivmax, double f,
                                          const double* g, double* req) {
    enum { L110, L135, L170, LEVAL, L240, L270, L280, LEND };
    int at;
    switch (s.pc) {
        case kVaFresh: {
            for (int i = 0; i < n; ++i) {
                V.v(kVaXaa, i) = V.v(kVaX, i);
                V.v(kVaGa, i) = 0.0;
                V.v(kVaD, i) = 0.0;
            }
            s.nfun = 0;
            s.itr = 0;
            s.c = 1.0;
            for (int k = 0; k < n * (n + 1) / 2; ++k) V.h(k) = 0.0;
            for (int i = 0, k = 0; i < n; ++i) {
                V.h(k) = 0.01 * s.c;
                k += n - i;
            }
            s.dff = 0.0;
            s.pc = kVaAtStart;
            for (int i = 0; i < n; ++i) req[i] = V.v(kVaXaa, i);
            return true;
        }
        case kVaAtStart:
            s.gesa = f;
            for (int i = 0; i < n; ++i) V.v(kVaG, i) = g[i];
            at = L110;
            break;
        case kVaInLine:
            s.fb = f;
            for (int i = 0; i < n; ++i) V.v(kVaGb, i) = g[i];
            at = LEVAL;
            break;
        default:
            return false;
    }
    for (;;) {
        switch (at) {
            case L110:
                s.fa = s.gesa;
                s.isfv = 1;
                for (int i = 0; i < n; ++i) V.v(kVaGa, i) = V.v(kVaG, i);
                at = L135;
                break;
            case L135: {
                ++s.itr;
                for (int i = 0; i < n; ++i) V.v(kVaD, i) = -V.v(kVaGa, i);
                va_solve(V, n, kVaD);
                s.c = 0.0;
                s.dga = 0.0;
                for (int i = 0; i < n; ++i) {
                    s.c = va_max(s.c, fabs(V.v(kVaD, i)));
                    s.dga = s.dga + V.v(kVaGa, i) * V.v(kVaD, i);
                }
                if (s.dga >= 0.0) {
                    at = L240;
                    break;
                }
                s.stmin = 0.0;
                s.stepbd = 0.0;
                s.steplb = 1e-9 / s.c;
                s.fmn = s.fa;
                s.gmin = s.dga;
                s.step = 1.0;
                if (s.dff <= 0.0) s.step = va_min(s.step, 1.0 / s.c);
                if (s.dff > 0.0) s.step = va_min(s.step, (s.dff + s.dff) / (-s.dga));
                at = L170;
                break;
            }
            case L170:
                s.c = s.stmin + s.step;
                if (s.nfun >= ivmax) {
                    at = LEND;
                    break;
                }
                ++s.nfun;
                for (int i = 0; i < n; ++i) {
                    V.v(kVaXbb, i) = V.v(kVaXaa, i) + s.c * V.v(kVaD, i);
                    req[i] = V.v(kVaXbb, i);
                }
                s.pc = kVaInLine;
                return true;
            case LEVAL: {
                if (s.isfv > 2) s.isfv = 2;
                bool better = s.fb < s.gesa;
                if (!better && !(s.fb > s.gesa)) {
                    double g1 = 0.0, g2 = 0.0;
                    for (int i = 0; i < n; ++i) {
                        g1 = g1 + V.v(kVaG, i) * V.v(kVaG, i);
                        g2 = g2 + V.v(kVaGb, i) * V.v(kVaGb, i);
                    }
                    better = g2 < g1;
                }
                if (better) {
                    s.isfv = 3;
                    s.gesa = s.fb;
                    for (int i = 0; i < n; ++i) {
                        V.v(kVaX, i) = V.v(kVaXbb, i);
                        V.v(kVaG, i) = V.v(kVaGb, i);
                    }
                }
                s.dgb = 0.0;
                for (int i = 0; i < n; ++i) s.dgb = s.dgb + V.v(kVaGb, i) * V.v(kVaD, i);
                if (s.fb - s.fa <= 0.1 * s.c * s.dga) at = L280;
                else if (s.step > s.steplb) at = L270;
                else at = L240;
                break;
            }
            case L240:
                at = s.isfv >= 2 ? L110 : LEND;
                break;
            case L270: {
                // cubic interpolation of the step
                s.stepbd = s.step;
                double cc = s.gmin + s.dgb - 3.0 * (s.fb - s.fmn) / s.step;
                cc = s.gmin / (cc + s.gmin - sqrt(cc * cc - s.gmin * s.dgb));
                s.c = cc;
                s.step = s.step * va_max(0.1, cc);
                at = L170;
                break;
            }
            case L280: {
                s.stepbd = s.stepbd - s.step;
                s.stmin = s.c;
                s.fmn = s.fb;
                s.gmin = s.dgb;
                s.step = 9.0 * s.stmin;
                if (s.stepbd > 0.0) s.step = 0.5 * s.stepbd;
                s.c = s.dga + 3.0 * s.dgb - 4.0 * (s.fb - s.fa) / s.stmin;
                if (s.c > 0.0) s.step = va_min(s.step, s.stmin * va_max(1.0, -s.dgb / s.c));
                if (s.dgb < 0.7 * s.dga) {
                    at = L170;
                    break;
                }
                s.isfv = 4 - s.isfv;
                if (s.stmin + s.step <= s.steplb) {
                    at = L240;
                    break;
                }
                s.ir = -n;
                for (int i = 0; i < n; ++i) {
                    V.v(kVaXaa, i) = V.v(kVaXbb, i);
                    V.v(kVaXbb, i) = V.v(kVaGa, i);
                    V.v(kVaD, i) = V.v(kVaGb, i) - V.v(kVaGa, i);
                    V.v(kVaGa, i) = V.v(kVaGb, i);
                }
                va_update(V, n, kVaXbb, 1.0 / s.dga, kVaW, s.ir, 1, 0.0);
                s.ir = -s.ir;
                va_update(V, n, kVaD, 1.0 / (s.stmin * (s.dgb - s.dga)), kVaD, s.ir, 0, 0.0);
                if (s.ir < n) {
                    at = LEND;
                    break;
                }
                s.dff = s.fa - s.fb;
                s.fa = s.fb;
                at = L135;
                break;
            }
            default:  // LEND
                s.pc = kVaFinished;
                return false;
        }
    }
}

}  // namespace att

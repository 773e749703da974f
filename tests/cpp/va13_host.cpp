// Test driver (not part of the library): the device minimizer's VA13 state
// machine (va13.h) run on the host with a callback objective, so it can be
// compared step for step with the oracle's restatement.
#include <vector>

#include "va13.h"

typedef void (*fg_t)(void* ctx, const double* x, double* f, double* g);

extern "C" int va13h_minimize(int n, const double* x0, int ivmax, fg_t fg, void* ctx, double* xlast,
                              double* flast, double* fbest, int* nfun) {
    if (n < 1 || n > att::kVaMaxDim) return -1;
    std::vector<double> st(att::va_elements(n), 0.0);
    const att::VaView V{st.data(), 1, n};
    for (int i = 0; i < n; ++i) V.v(att::kVaX, i) = x0[i];
    att::VaState s{};
    s.pc = att::kVaFresh;
    double req[att::kVaMaxDim], g[att::kVaMaxDim], f = 0.0;
    bool more = att::va_resume(s, V, n, ivmax, 0.0, nullptr, req);
    while (more) {
        fg(ctx, req, &f, g);
        for (int i = 0; i < n; ++i) xlast[i] = req[i];
        *flast = f;
        more = att::va_resume(s, V, n, ivmax, f, g, req);
    }
    *fbest = s.gesa;
    *nfun = s.nfun;
    return 0;
}

// The factor update alone (packed factors a, vectors z and w of length n;
// alias != 0 passes z as w too, as minfor's second update does).
extern "C" void va13h_update(double* a, int n, double* z, double sig, double* w, int* ir, int mk,
                             double eps, int alias) {
    std::vector<double> st(att::va_elements(n), 0.0);
    const att::VaView V{st.data(), 1, n};
    for (int i = 0; i < n; ++i) {
        V.v(att::kVaXbb, i) = z[i];
        V.v(att::kVaW, i) = w[i];
    }
    for (int k = 0; k < n * (n + 1) / 2; ++k) V.h(k) = a[k];
    att::va_update(V, n, att::kVaXbb, sig, alias ? att::kVaXbb : att::kVaW, *ir, mk, eps);
    for (int i = 0; i < n; ++i) {
        z[i] = V.v(att::kVaXbb, i);
        w[i] = alias ? z[i] : V.v(att::kVaW, i);
    }
    for (int k = 0; k < n * (n + 1) / 2; ++k) a[k] = V.h(k);
}

// gpk_host.hpp — host-side helpers shared by the library's translation units:
// error type and status mapping, device buffers, and the weight tables of the
// Gaussian geometry, projection and remap stencils.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpk_c.h"
#include "gpk_cell.hpp"

namespace gpkh {

using gpkd::RowLutDesc;

// thread-local message behind gpk_last_error()
std::string& last_error();


struct GpkError : std::runtime_error {
    int code;
    GpkError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw GpkError(GPK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int nblk(long long n, int b) { return int((n + b - 1) / b); }
inline int gsblk(long long n, int b) { return std::max(1, std::min(nblk(n, b), 8192)); }

// ---------------------------------------------------------------------------
// Device buffer helper
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void swap(DevBuf& o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
    void alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            throw GpkError(GPK_E_OOM, "hipMalloc(" + std::to_string(b) + ") failed");
        }
        bytes = b;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
    ~DevBuf() { release(); }
};

// ---------------------------------------------------------------------------
// Host-side weight construction (include/gpk_core.hpp:37-55,
// src/gpk_geometry.cpp:12-36).  The float build uses the same float
// expressions as the reference so the tables are bit-identical.
// ---------------------------------------------------------------------------
struct Weights {
    int radius = 0;        // Gaussian radius
    int rg = 0;            // geometry stencil radius
    float sigma_cells = 0.f;
    std::vector<double> w0, w1, w2;   // stored as double, exact for float values
};

inline Weights gauss_float(float sigma) {
    Weights W;
    W.sigma_cells = std::max(sigma, 0.5f);
    const float s = W.sigma_cells;
    W.radius = static_cast<int>(std::ceil(3.5f * s));
    const int n = 2 * W.radius + 1;
    std::vector<float> w(n);
    float sum = 0.f;
    for (int i = 0; i < n; ++i) {
        const float x = float(i - W.radius);
        w[i] = std::exp(-x * x / (2.f * s * s));
        sum += w[i];
    }
    for (auto& v : w) v /= sum;
    W.rg = W.radius + 2;
    const int width = 2 * W.rg + 1;
    std::vector<float> a(width, 0.f), b(width, 0.f), c(width, 0.f);
    const int o1[4] = {-2, -1, 1, 2};
    const float c1[4] = {1.f / 12.f, -8.f / 12.f, 8.f / 12.f, -1.f / 12.f};
    const int o2[5] = {-2, -1, 0, 1, 2};
    const float c2[5] = {-1.f / 12.f, 16.f / 12.f, -30.f / 12.f, 16.f / 12.f, -1.f / 12.f};
    auto at = [&](int o) { return (o < -W.radius || o > W.radius) ? 0.f : w[o + W.radius]; };
    for (int o = -W.rg; o <= W.rg; ++o) {
        const int j = o + W.rg;
        a[j] = at(o);
        for (int k = 0; k < 4; ++k) b[j] += c1[k] * at(o - o1[k]);
        for (int k = 0; k < 5; ++k) c[j] += c2[k] * at(o - o2[k]);
    }
    W.w0.assign(a.begin(), a.end());
    W.w1.assign(b.begin(), b.end());
    W.w2.assign(c.begin(), c.end());
    return W;
}

inline Weights gauss_double(float sigma) {
    Weights W;
    W.sigma_cells = std::max(sigma, 0.5f);
    const double s = W.sigma_cells;
    W.radius = static_cast<int>(std::ceil(3.5f * W.sigma_cells));
    const int n = 2 * W.radius + 1;
    std::vector<double> w(n);
    double sum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double x = double(i - W.radius);
        w[i] = std::exp(-x * x / (2.0 * s * s));
        sum += w[i];
    }
    for (auto& v : w) v /= sum;
    W.rg = W.radius + 2;
    const int width = 2 * W.rg + 1;
    W.w0.assign(width, 0.0); W.w1.assign(width, 0.0); W.w2.assign(width, 0.0);
    const int o1[4] = {-2, -1, 1, 2};
    const double c1[4] = {1.0 / 12.0, -8.0 / 12.0, 8.0 / 12.0, -1.0 / 12.0};
    const int o2[5] = {-2, -1, 0, 1, 2};
    const double c2[5] = {-1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0};
    auto at = [&](int o) { return (o < -W.radius || o > W.radius) ? 0.0 : w[o + W.radius]; };
    for (int o = -W.rg; o <= W.rg; ++o) {
        const int j = o + W.rg;
        W.w0[j] = at(o);
        for (int k = 0; k < 4; ++k) W.w1[j] += c1[k] * at(o - o1[k]);
        for (int k = 0; k < 5; ++k) W.w2[j] += c2[k] * at(o - o2[k]);
    }
    return W;
}

// exp(-r^2/(2 s^2)) indexed by integer r^2 (src/gpk_physics.cpp:160-185,
// src/gpk_transitions.cpp:50-81): expf promoted to double in float modes.
inline std::vector<double> radial_table(float s, int r, bool single) {
    const int n = 3 * r * r + 1;
    std::vector<double> t(n);
    if (single) {
        const float inv = 0.5f / (s * s);
        for (int r2 = 0; r2 < n; ++r2) {
            const float r2f = static_cast<float>(r2);
            t[r2] = static_cast<double>(std::exp(-r2f * inv));
        }
    } else {
        const double sd = s;
        const double inv = 0.5 / (sd * sd);
        for (int r2 = 0; r2 < n; ++r2) t[r2] = std::exp(-double(r2) * inv);
    }
    return t;
}

// Row-sum tables of the separable modes (same construction as the device
// reads them, see RowLutDesc): nch = ceil(W/9) chunks of C = ceil(W/nch) dx,
// entry base_k + idx*3 + a = sum of weight set a over the set bits of idx in
// ascending dx (bit j of idx is the chunk's j-th dx).
template <class R>
RowLutDesc make_row_lut(const std::vector<R>& w0, const std::vector<R>& w1, const std::vector<R>& w2,
                        std::vector<R>& t) {
    RowLutDesc L{};
    L.W = static_cast<int>(w0.size());
    L.nch = (L.W + 8) / 9;
    const int C = (L.W + L.nch - 1) / L.nch;
    int off = 0;
    for (int k = 0; k < L.nch; ++k) {
        L.start[k] = k * C;
        L.len[k] = std::min(C, L.W - L.start[k]);
        L.base[k] = off;
        off += (1 << L.len[k]) * 3;
    }
    t.assign(off, R(0));
    const std::vector<R>* w[3] = {&w0, &w1, &w2};
    for (int k = 0; k < L.nch; ++k)
        for (int idx = 0; idx < (1 << L.len[k]); ++idx)
            for (int a = 0; a < 3; ++a) {
                R sum = R(0);
                for (int j = 0; j < L.len[k]; ++j)
                    if ((idx >> j) & 1) sum += (*w[a])[L.start[k] + j];
                t[L.base[k] + idx * 3 + a] = sum;
            }
    return L;
}

// Product table of the direct masked stencil: entry t*C + k for window
// position t (dz, dy, dx order) holds (gx*gy)*gz of component k, C = 9 in 3-D
// (Ix, Iy, Ixx, Iyy, Ixy, Iz, Izz, Ixz, Iyz) and 5 in 2-D (gz = 1), the
// reference's product order (src/gpk_geometry.cpp:112-133).
template <class R>
std::vector<R> direct_table(const std::vector<R>& w0, const std::vector<R>& w1, const std::vector<R>& w2,
                            bool dim3) {
    const int Wd = static_cast<int>(w0.size());
    std::vector<R> t;
    if (dim3) {
        t.resize(size_t(Wd) * Wd * Wd * 9);
        size_t o = 0;
        for (int jz = 0; jz < Wd; ++jz)
            for (int jy = 0; jy < Wd; ++jy)
                for (int jx = 0; jx < Wd; ++jx) {
                    const R gx0 = w0[jx], gx1 = w1[jx], gx2 = w2[jx];
                    const R gy0 = w0[jy], gy1 = w1[jy], gy2 = w2[jy];
                    const R gz0 = w0[jz], gz1 = w1[jz], gz2 = w2[jz];
                    t[o++] = gx1 * gy0 * gz0; t[o++] = gx0 * gy1 * gz0;
                    t[o++] = gx2 * gy0 * gz0; t[o++] = gx0 * gy2 * gz0;
                    t[o++] = gx1 * gy1 * gz0; t[o++] = gx0 * gy0 * gz1;
                    t[o++] = gx0 * gy0 * gz2; t[o++] = gx1 * gy0 * gz1;
                    t[o++] = gx0 * gy1 * gz1;
                }
    } else {
        t.resize(size_t(Wd) * Wd * 5);
        size_t o = 0;
        const R one = R(1);
        for (int jy = 0; jy < Wd; ++jy)
            for (int jx = 0; jx < Wd; ++jx) {
                const R gx0 = w0[jx], gx1 = w1[jx], gx2 = w2[jx];
                const R gy0 = w0[jy], gy1 = w1[jy], gy2 = w2[jy];
                t[o++] = gx1 * gy0 * one; t[o++] = gx0 * gy1 * one;
                t[o++] = gx2 * gy0 * one; t[o++] = gx0 * gy2 * one;
                t[o++] = gx1 * gy1 * one;
            }
    }
    return t;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        last_error().clear();
        return GPK_OK;
    } catch (const GpkError& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return GPK_E_INVALID;
    } catch (...) {
        last_error() = "unknown error";
        return GPK_E_INVALID;
    }
}

} // namespace gpkh
